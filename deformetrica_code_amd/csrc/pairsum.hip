// pairsum.hip — host side of the pair-sum engine: per-thread error and timing state,
// the Hilbert ordering of the targets (bounding box, 30-bit keys, rocPRIM radix sort),
// the work plan and workspace layout of a call, and the launch of the per-precision
// kernel instantiations (pairsum_f64.hip, pairsum_f32c.hip, pairsum_f32x.hip).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "dfm_hip.h"
#include "dfm_internal.h"
#include "pairsum_kernels.h"

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
  double pairs = 0.0, flops = 0.0;  // algorithmic work of the timed launches
};
static thread_local ProfileState g_profile;

void profile_enable(bool on) { g_profile.on = on; }

// Timing of a launch made outside the pair-sum entry points (the persistent shoot): an
// event before it and, with its algorithmic work, one after it, while profiling is on.
hipEvent_t profile_begin(hipStream_t st) {
  if (!g_profile.on) return nullptr;
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  (void)hipEventRecord(e, st);
  return e;
}

void profile_end(hipStream_t st, hipEvent_t e0, double pairs, double flops) {
  if (!e0) return;
  hipEvent_t e1 = nullptr;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return;
  }
  (void)hipEventRecord(e1, st);
  g_profile.events.push_back({e0, e1});
  g_profile.pairs += pairs;
  g_profile.flops += flops;
}

int profile_work(double *pairs, double *flops) {
  if (pairs) *pairs = g_profile.pairs;
  if (flops) *flops = g_profile.flops;
  g_profile.pairs = g_profile.flops = 0.0;
  return DFM_OK;
}

// Algorithmic flops of one (target, source) pair of an op, counted on the reference's
// per-pair arithmetic (ExactKernel.cxx, AbstractKernel.cxx; exp counted as one): the
// kernel value 3D + 1 (D differences, D squares and D - 1 adds, the scale, the exp), then
//   Convolve        2 Dv (multiply-add per weight)
//   Grad1           2 Dv - 1 (alpha . w) + 2 (coefficient) + 2 D
//   varlin          1 + D (coefficient times difference) + 2 Dv D
//   gradient(dim)   1 + 2 Dv
//   Hessian         D(D+1)/2 (d d^T) + Dv (2 + D(D+1))
//   SpecialHessian  2 Dv - 1 (beta . beta) + D + 2D - 1 (eta difference and dot) + 2 + 4 D
//   shooting VJP    4D - 1 + 2D - 1 + D + 2D - 1 + 4 + 10 D  (see pair_update)
double flops_per_pair(int op, int D, int DV) {
  const double k = 3.0 * D + 1.0;
  const double conv = 2.0 * DV, grad = 2.0 * DV - 1.0 + 2.0 + 2.0 * D;
  switch (op) {
    case OP_CONV: return k + conv;
    case OP_GRAD1: return k + grad;
    case OP_CONVGRAD1: return k + conv + grad;
    case OP_VARLIN: return k + 1.0 + D + 2.0 * DV * D;
    case OP_GRADDIM: return k + 1.0 + 2.0 * DV;
    case OP_HESS: return k + D * (D + 1) / 2.0 + DV * (2.0 + D * (D + 1));
    case OP_GRADDIFF: return k + 2.0 * DV - 1.0 + 3.0 * D - 1.0 + 2.0 + 4.0 * D;
    case OP_SHOOTVJP: return k + 19.0 * D + 1.0;
    case OP_VARIFOLD: return k + conv + grad + 1.0;  // the fused sum, Convolve weight K (u.u)
    case OP_VARIFOLD_CONV: return k + conv + 2.0 * DV;  // its Convolve part: u.u, K (u.u)
  }
  return k;
}

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
// Hilbert ordering of the targets (FP32C, BF16MFMA)

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

// Hilbert index of a cell with N-dimensional coordinates X[] of B bits each: J. Skilling,
// "Programming the Hilbert curve", AIP Conf. Proc. 707, 381 (2004): the axes are
// transformed in place into the "transpose" of the index (undo the excess work of each
// level's inversions and exchanges, then Gray-encode), whose bits interleaved, most
// significant first and X[0] first within a bit, give the index.  Consecutive indices are
// face-adjacent cells, so runs of sorted targets stay compact where a Morton (Z) order
// jumps across octant boundaries.
template <int N, int B>
__device__ __forceinline__ void hilbert_transpose(uint32_t (&X)[N]) {
  for (uint32_t Q = 1u << (B - 1); Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
#pragma unroll
    for (int i = 0; i < N; i++) {
      if (X[i] & Q) {
        X[0] ^= P;  // invert the low bits of X[0]
      } else {      // exchange the low bits of X[0] and X[i]
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
#pragma unroll
  for (int i = 1; i < N; i++) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = 1u << (B - 1); Q > 1; Q >>= 1)
    if (X[N - 1] & Q) t ^= Q - 1;
#pragma unroll
  for (int i = 0; i < N; i++) X[i] ^= t;
}

__global__ __launch_bounds__(256) void k_curve_keys(const double *x, long n, int D,
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
  if (D == 3) {
    uint32_t X[3] = {q[0], q[1], q[2]};
    hilbert_transpose<3, 10>(X);
    keys[i] = (spread3(X[0]) << 2) | (spread3(X[1]) << 1) | spread3(X[2]);
  } else {
    uint32_t X[2] = {q[0], q[1]};
    hilbert_transpose<2, 15>(X);
    keys[i] = (spread2(X[0]) << 1) | spread2(X[1]);
  }
  vals[i] = (uint32_t)i;
}

// Device-to-device copy of 16-byte groups (plus a byte tail); also the known-size
// reference kernel for scaling HBM traffic counters.
__global__ __launch_bounds__(256) void k_copy(uint4 *dst, const uint4 *src, long n16,
                                              unsigned char *dtail, const unsigned char *stail,
                                              int ntail) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) dtail[threadIdx.x] = stail[threadIdx.x];
}

// The fused integrator step (KParams e_*) applied from the stored sums out0 (conv) and
// out1 (grad), for the kernels that do not apply it themselves; segment blockIdx.y.
template <int D>
__global__ __launch_bounds__(256) void k_step_update(KParams p) {
  const int seg = blockIdx.y;
  long lo = 0, hi = p.nx;
  if (p.xoff) {
    lo = p.xoff[seg];
    hi = p.xoff[seg + 1];
  }
  const long o = lo + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= hi) return;
  double conv[D], grad[D];
#pragma unroll
  for (int k = 0; k < D; k++) {
    conv[k] = p.out0[o * D + k];
    grad[k] = p.out1[o * D + k];
  }
  step_update<D>(p, o, seg, conv, grad);
}

struct LongChunk {
  long v[32];
};

__global__ __launch_bounds__(64) void k_put_longs(long *dst, LongChunk c, int n) {
  if ((int)threadIdx.x < n) dst[threadIdx.x] = c.v[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_fill(double *dst, long n, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

// Empty sums (S = 0) of a call with the linear epilogue (KParams o_*) on: out0[i < n0] and
// out1[i < n1] take the epilogue's other terms.
__global__ __launch_bounds__(256) void k_put_empty(KParams p, long n0, long n1) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p.out0 && i < n0) put<0>(p, p.out0, i, 0.0);
  if (p.out1 && i < n1) put<1>(p, p.out1, i, 0.0);
}

// ------------------------------------------------------------------------------------
// planning

static int targets_per_lane(int op, int prec, int D, int DV) {
  if (prec == DFM_FP64) return 1;
  switch (op) {
    case OP_VARLIN: return DV * D > 9 ? 1 : 2;
    case OP_GRADDIFF: return 2;
    case OP_SHOOTVJP: return 1;
    case OP_HESS: return 1;
    default: return 4;
  }
}

static int tile_of(int prec) {
  return prec == DFM_FP64 ? kTileF64 : prec == DFM_BF16MFMA ? kBfTile : kTileF32;
}

// DFM_BF16MFMA runs the matrix-core kernel for Convolve / Grad1 / fused sums with at least
// kBfMinSources sources (in the largest segment): single calls large enough to be
// Hilbert-sorted, and ragged batches (nseg > 1, the caller's target order: groups of 256
// targets that span too far take the FP32C code inside the kernel); everything else runs
// as DFM_FP32C.
// 64 segments of the 10,242 icosphere vertices (in the order of first use by the faces)
// against control grids of 2,048, 4,096, 8,192 and 16,384 points (sigma 0.25), with splits
// of at least four source tiles, ran the fused sum 0.84, 0.84, 0.86 and 0.81 times as long
// as DFM_FP32C (Convolve 0.82, 0.75, 0.71, 0.71; Grad1 0.87, 0.83, 0.84, 0.83).  With
// one-tile splits the same batches took 1.02-1.16 times as long up to 4,096 sources: each
// group's set-up (the target splits, the FP64 epilogue) and its split partials pay off only
// over several source tiles.
constexpr long kBfMinSources = 2048;
// Ragged batches also need this many targets in their largest segment.  The 64-subject
// atlas's sums over its 2,048 control points per subject (a 16 x 16 x 8 grid over
// [-1.5, 1.5]^3 at sigma 0.25: every group of 256 targets, and every wavefront, spans too far
// for the matrix-core path) took 169 us per fused shooting step and 733 us per flow pull
// (onto 10,242 template points) in the FP32C code inside the matrix-core kernel, against 141
// and 653 us as DFM_FP32C with two targets per lane.
constexpr long kBfMinBatchTargets = 4096;

static int effective_prec(int op, int prec, int D, int DV, int nseg, long nx, long ny) {
  if (prec != DFM_BF16MFMA) return prec;
  const bool big =
      ny >= kBfMinSources && (nseg > 1 ? nx >= kBfMinBatchTargets : nx >= kSortMinTargets);
  return (big && bf_supported(op, D, DV)) ? prec : DFM_FP32C;
}

static int nacc_of(int op, int D, int DV) {
  switch (op) {
    case OP_CONV: return DV;
    case OP_GRAD1: return D;
    case OP_CONVGRAD1: return DV + D;
    case OP_VARIFOLD: return DV + D;
    case OP_VARIFOLD_CONV: return DV;
    case OP_VARLIN: return DV * D;
    case OP_GRADDIM: return DV;
    case OP_GRADDIFF: return 2 * D;
    case OP_HESS: return DV * (1 + D * (D + 1) / 2);
    default: return 5 * D;
  }
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// The tunable part of a plan: targets per lane, the wavefront count the source splits
// aim for, and the fewest source tiles a split may hold.
struct Choice {
  int R, waves, min_tiles;
  int mx = 0;  // FP32C fused Convolve + Grad1 on the matrix cores (R = 1)
  int w = 1;   // wavefronts per workgroup on the same targets (kWaveSplit: FP32, R = 1)
};

// FP32 grids of fewer target workgroups than this may run k_pairsum's small-grid form
// (autotuner candidates): kWaveSplit wavefronts to a workgroup on the same targets, so the
// source splits are added in LDS instead of in split partials.  It pays where it allows
// more wavefronts than split partials would: the 4,096-point registration's fused shooting
// step (64 target workgroups) takes 0.025 ms with 16 one-tile splits per workgroup and 4
// workgroup splits against 0.033 ms with 16 four-tile splits; with 160 or 320 workgroups
// (its flow and data-term sums) it lost 3-25% to the one-wavefront form, whose 1,024-lane
// workgroups fill fewer CUs.
constexpr long kWaveSplitGrid = 1024;

// the ops of the small-grid form: those of the expanded-distance form with up to 3 weights
// (one target per lane fits the 128 VGPRs of a 1,024-lane workgroup without spills), and
// SHOOTVJP in 512-lane workgroups (wave_split_of).  The shooting adjoint's SHOOTVJP sum of
// the 4,096-point registration (64 target workgroups) took 28 us plus a 10 us split
// reduction this way against 43 plus 16 us in the one-wavefront form.
static bool wave_split_op(int op, int DV) {
  return DV <= 3 && (op == OP_CONV || op == OP_GRAD1 || op == OP_CONVGRAD1 ||
                     op == OP_VARIFOLD || op == OP_VARIFOLD_CONV || op == OP_SHOOTVJP);
}

// Heuristic, from the autotuner's measurements on the shapes of the workloads: one target
// per lane (64-target wavefronts: the fewest registers, the most wavefronts per SIMD and
// the tightest wavefront boxes); source splits aiming for 262144 wavefronts on grids of
// >= 4096 target workgroups and 16384 below (the short tail of many small wavefronts
// balances the CUs); splits of at least 4 source tiles, so that the split reduction
// stays short, except on the smallest grids (one tile: FP32 below 64 target workgroups,
// FP64 below 1024).  At N = M = 1M the fused FP32C sum runs in 363 ms with this
// plan (366 ms aiming for 65536 wavefronts, same box) against 483 ms with 4 targets per
// lane, 4096 wavefronts and 256-source tiles.
static Choice default_choice(int op, int prec, int D, int DV, int nseg, long max_nx) {
  (void)D; (void)DV;
  // DFM_BF16MFMA: workgroups of 256 targets (R = 4 in 64-target units) and 256-source tiles;
  // the split count aims for 65536 workgroups.  At N = M = 1M (17 splits) the fused sum runs
  // in 264 ms against 280 ms aiming for 16384 (4 splits) and 303 ms for 4096, same box,
  // every candidate timed by the autotuner: the shorter splits also spread the FP32C
  // wavefronts of the few wide groups over more of the chip.
  // Splits of at least four tiles (see kBfMinSources; the 1M-source headline takes 17
  // splits of 230 tiles either way); ragged batches aim for 16,384 workgroups (the
  // 64-subject atlas's data-term sums take 9.4-9.7 ms either way).
  if (prec == DFM_BF16MFMA) return nseg > 1 ? Choice{4, 16384, 4} : Choice{4, 65536, 4};
  const long blocks = (max_nx + 63) / 64 * nseg;
  // FP32 Convolve / Grad1 / fused sums (the expanded-distance form) on large single grids:
  // two targets per lane share each LDS source record, halving the LDS reads that bound
  // the one-target plan.  At N = M = 1M the fused FP32C sum runs in 318 ms with two
  // targets per lane and 65536 wavefronts against 354 ms with the one-target plan below
  // (same box, every candidate timed by the autotuner).
  const bool expandable =
      op == OP_CONV || op == OP_GRAD1 || op == OP_CONVGRAD1 || op == OP_VARIFOLD ||
      op == OP_VARIFOLD_CONV;
  // Grad1 and the fused sum with more than 3 weights keep one target per lane (the second
  // target's accumulators cost occupancy): a 1M-sample fused Dv = 6 sum runs 446 ms with
  // R = 1 and 262,144 wavefronts against 481 ms with R = 2, a 64-subject batch of them
  // 11.8 ms with R = 1 and 65,536 wavefronts against 15.9 with R = 2.  So does the
  // varifold's fused sum (the per-pair u_i . u_j on top of six accumulators): the
  // 64-subject atlas's two data-term batches take 9.6 and 10.0 ms with R = 1 and 65,536
  // wavefronts against 12.9 and 14.7 ms with R = 2.
  const bool two =
      expandable && (op == OP_CONV || op == OP_VARIFOLD_CONV || (DV <= 3 && op != OP_VARIFOLD));
  if (prec != DFM_FP64 && expandable && nseg == 1 && blocks >= 4096)
    return two ? Choice{2, 65536, 4} : Choice{1, 262144, 4};
  // ragged batches on grids of >= 1,024 target workgroups (the atlas's flow, adjoint and
  // data-term sums): the 64-subject atlas's batched Convolve over the template vertices
  // runs 0.431 ms with R = 2 and 16,384 wavefronts against 0.525 ms with the plan below
  if (prec != DFM_FP64 && expandable && nseg > 1 && blocks >= 1024)
    return two ? Choice{2, 16384, 4} : Choice{1, 65536, 4};
  // at 1,000 x 1,000 (16 target workgroups) the fused FP32C sum takes 0.017 ms with
  // one-tile splits against 0.030 ms with four-tile ones (FP64: 0.021 against 0.053 ms);
  // from 64 workgroups up (a 4,096-point registration's sums) one-tile FP32 splits cost
  // more in split partials than they gain: 16.4 against 11.8 ms for that gradient
  // the small-grid form where it pays (see kWaveSplitGrid), aiming for 256 workgroups: the
  // registration's 4,096 x 8,192 fused sum takes 0.040 ms so against 0.054 ms aiming for
  // 1,024 workgroups of one-tile splits
  if (prec != DFM_FP64 && wave_split_op(op, DV) && blocks >= 32 && blocks < 128)
    return Choice{1, 4096, 1, 0, kWaveSplit};
  const int min_tiles = (blocks < 64 || (prec == DFM_FP64 && blocks < 1024)) ? 1 : 4;
  return Choice{1, blocks >= 4096 ? 262144 : 16384, min_tiles};
}

// Candidates of the autotuner, the default first.  R never exceeds the op's default (the
// register budget of the accumulators).
static std::vector<Choice> candidates(int op, int prec, int D, int DV, int nseg, long max_nx) {
  const Choice d = default_choice(op, prec, D, DV, nseg, max_nx);
  if (prec == DFM_BF16MFMA) {
    std::vector<Choice> out{d};
    for (int waves : {4096, 16384, 65536})
      for (int mt : {1, 2})
        if (!(waves == d.waves && mt == d.min_tiles)) out.push_back(Choice{4, waves, mt});
    return out;
  }
  const int Rmax = targets_per_lane(op, prec, D, DV);
  std::vector<int> Rs{Rmax};
  if (Rmax >= 4) Rs.push_back(2);
  if (Rmax >= 2) Rs.push_back(1);
  std::vector<Choice> out{d};
  auto same = [](const Choice &a, const Choice &b) {
    return a.R == b.R && a.waves == b.waves && a.min_tiles == b.min_tiles && a.mx == b.mx &&
           a.w == b.w;
  };
  // long grids: finer splits shorten the tail further (one target per lane only, so the
  // split partials of the largest candidate, and the workspace, do not grow)
  const long blocks = (max_nx + 63) / 64 * nseg;
  for (int R : Rs)
    for (int waves : {4096, 16384, 65536, 131072, 262144})
      for (int mt : {4, 1}) {
        if (waves > 65536 && (R != 1 || blocks < 4096)) continue;
        Choice c{R, waves, mt, 0};
        if (!same(c, d)) out.push_back(c);
      }
  if (prec == DFM_FP32C && mx_supported(op, D, DV))
    for (int waves : {4096, 16384})
      for (int mt : {4, 1}) {
        Choice c{1, waves, mt, 1};
        if (!same(c, d)) out.push_back(c);
      }
  if (prec != DFM_FP64 && blocks < kWaveSplitGrid && wave_split_op(op, DV))
    for (int waves : {4096, 16384, 65536})
      for (int mt : {4, 1}) {
        Choice c{1, waves, mt, 0, kWaveSplit};
        if (!same(c, d)) out.push_back(c);
      }
  return out;
}

// Work split shared by single calls and ragged batches: chunks of the largest segment's
// targets, and source splits of the largest segment's sources chosen so that the grid
// holds about c.waves wavefronts with >= c.min_tiles source tiles per split.
static void split_plan(Plan &pl, int op, int prec, int D, int DV, int nseg, long max_nx,
                       long max_ny, const Choice &c) {
  (void)op; (void)D; (void)DV;
  pl.R = c.mx ? 1 : c.R;
  pl.mx = c.mx;
  pl.W = (c.w > 1 && !c.mx && pl.R == 1 && prec != DFM_FP64 && wave_split_op(op, DV))
             ? wave_split_of(op)
             : 1;
  pl.nseg = nseg;
  pl.chunks = (max_nx + 64L * pl.R - 1) / (64L * pl.R);
  const int TS = tile_of(prec);
  const long tiles = (max_ny + TS - 1) / TS;
  const long blocks = pl.chunks * nseg;
  long want = blocks > 0 ? (c.waves + blocks - 1) / blocks : 1;
  long maxsplit = std::max(1L, tiles / c.min_tiles);
  long ns = std::max(1L, std::min(want, maxsplit));
  ns = std::min(ns, 65535L);
  if (pl.W > 1) {
    // ns wavefront splits, W to a workgroup: ceil(ns / W) workgroup splits of W * split_len
    const long nsw = (ns + pl.W - 1) / pl.W;
    const long per = (tiles + nsw * pl.W - 1) / (nsw * pl.W);
    pl.split_len = std::max(1L, per) * TS;
    const long span = pl.split_len * pl.W;
    pl.nsplit = (int)std::max(1L, (max_ny + span - 1) / span);
    return;
  }
  const long per = (tiles + ns - 1) / ns;
  pl.split_len = std::max(1L, per) * TS;
  pl.nsplit = (int)std::max(1L, (max_ny + pl.split_len - 1) / pl.split_len);
}

static Plan batch_plan(int op, int prec, int D, int DV, int nseg, long nx_total, long max_nx,
                       long max_ny, const Choice &c) {
  Plan pl{};
  split_plan(pl, op, prec, D, DV, nseg, max_nx, max_ny, c);
  pl.sort = false;
  size_t off = 0;
  if (pl.nsplit > 1) {
    pl.off_partial = off;
    off = align256(off + sizeof(double) * (size_t)pl.nsplit * nx_total * nacc_of(op, D, DV));
  }
  pl.work_bytes = off;
  return pl;
}

static Plan single_plan(int op, int prec, int D, int DV, long nx, long ny, const Choice &c) {
  Plan pl{};
  split_plan(pl, op, prec, D, DV, 1, nx, ny, c);
  pl.sort = prec != DFM_FP64 && nx >= kSortMinTargets;
  size_t off = 0;
  if (pl.sort) {
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

// Workspace bound over every candidate, so that a caller's buffer fits whatever plan the
// autotuner picks.
Plan make_plan(int op, int prec, int D, int DV, long nx, long ny) {
  prec = effective_prec(op, prec, D, DV, 1, nx, ny);
  Plan best = single_plan(op, prec, D, DV, nx, ny, default_choice(op, prec, D, DV, 1, nx));
  for (const Choice &c : candidates(op, prec, D, DV, 1, nx)) {
    Plan p = single_plan(op, prec, D, DV, nx, ny, c);
    if (p.work_bytes > best.work_bytes) best = p;
  }
  return best;
}

Plan make_batch_plan(int op, int prec, int D, int DV, int nseg, long nx_total, long max_nx,
                     long max_ny) {
  prec = effective_prec(op, prec, D, DV, nseg, max_nx, max_ny);
  Plan best = batch_plan(op, prec, D, DV, nseg, nx_total, max_nx, max_ny,
                         default_choice(op, prec, D, DV, nseg, max_nx));
  for (const Choice &c : candidates(op, prec, D, DV, nseg, max_nx)) {
    Plan p = batch_plan(op, prec, D, DV, nseg, nx_total, max_nx, max_ny, c);
    if (p.work_bytes > best.work_bytes) best = p;
  }
  return best;
}

// ------------------------------------------------------------------------------------
// autotuner: per (op, precision, d, dv, batched, log2 of the largest target and source
// counts, grid class), time every candidate twice on real calls (HIP events around the
// pair-sum launch, read back at a later call of the same key) and keep the fastest.
//
// The grid class (target workgroups of the whole grid: < 64, < 1024, < 4096, >= 4096) is
// part of the key because the candidate set and the default plan depend on it: a winner
// tuned on a large batch is then never reused by a smaller batch of the same largest
// segment, whose workspace (sized over its own candidates) would not hold it.

struct TuneKey {
  int op, prec, D, DV, batch, lx, ly, grid;
  bool operator<(const TuneKey &o) const {
    const int a[8] = {op, prec, D, DV, batch, lx, ly, grid};
    const int b[8] = {o.op, o.prec, o.D, o.DV, o.batch, o.lx, o.ly, o.grid};
    return std::lexicographical_compare(a, a + 8, b, b + 8);
  }
};

static int grid_class(int nseg, long max_nx) {
  const long blocks = (max_nx + 63) / 64 * nseg;
  return blocks < 64 ? 0 : blocks < 1024 ? 1 : blocks < 4096 ? 2 : 3;
}

// One timed launch of a candidate.  The events are created under the tuner's lock and
// shared between the launch that records them and the entry that reads them, so neither
// a concurrent call of the same key nor dfm_autotune_reset can free them under the other.
struct TimedLaunch {
  int cand = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::atomic<int> state{0};  // 0 being issued, 1 recorded, 2 launch failed
  ~TimedLaunch() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

struct TuneEntry {
  std::vector<Choice> cand;
  std::vector<float> best_ms;
  int next = 0;
  std::vector<std::shared_ptr<TimedLaunch>> pending;
  bool done = false;
  int winner = 0;
};

static std::mutex g_tune_mu;
static std::map<TuneKey, TuneEntry> g_tune;
static std::atomic<int> g_tune_mode{0};
constexpr int kTuneReps = 2;

static int ilog2(long v) {
  int l = 0;
  while (v > 1) {
    v >>= 1;
    l++;
  }
  return l;
}

void tune_set_mode(int mode) { g_tune_mode = mode; }

void tune_reset() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();  // launches still in flight keep their own TimedLaunch alive
}

// Read back every timed launch of the entry that has been recorded (caller holds the lock).
static void resolve_pending(TuneEntry &t) {
  std::vector<std::shared_ptr<TimedLaunch>> keep;
  for (auto &tl : t.pending) {
    const int st = tl->state.load();
    if (st == 0) {  // another thread is still issuing it
      keep.push_back(tl);
      continue;
    }
    float ms = 0.f;
    if (st == 1 && hipEventSynchronize(tl->e1) == hipSuccess &&
        hipEventElapsedTime(&ms, tl->e0, tl->e1) == hipSuccess)
      t.best_ms[tl->cand] = std::min(t.best_ms[tl->cand], ms);
  }
  t.pending.swap(keep);
  if (t.next >= (int)t.cand.size() * kTuneReps && t.pending.empty() && !t.done) {
    t.done = true;
    t.winner = (int)(std::min_element(t.best_ms.begin(), t.best_ms.end()) - t.best_ms.begin());
  }
}

// The plan choice for a call and, in explore mode, the timed launch to record.
static Choice tune_choose(int op, int prec, int D, int DV, int batch, int nseg, long max_nx,
                          long max_ny, std::shared_ptr<TimedLaunch> *timed) {
  timed->reset();
  const int mode = g_tune_mode;
  const Choice d = default_choice(op, prec, D, DV, nseg, max_nx);
  if (mode == 0) return d;
  const TuneKey k{op, prec, D, DV, batch, ilog2(max_nx), ilog2(max_ny), grid_class(nseg, max_nx)};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(k);
  if (it == g_tune.end()) {
    if (mode == 2) return d;
    TuneEntry t;
    t.cand = candidates(op, prec, D, DV, nseg, max_nx);
    t.best_ms.assign(t.cand.size(), 3.4e38f);
    it = g_tune.emplace(k, std::move(t)).first;
  }
  TuneEntry &t = it->second;
  if (mode == 1) resolve_pending(t);
  if (t.done) return t.cand[t.winner];
  if (mode == 2 || t.next >= (int)t.cand.size() * kTuneReps) return d;
  auto tl = std::make_shared<TimedLaunch>();
  tl->cand = t.next / kTuneReps;
  if (hipEventCreate(&tl->e0) != hipSuccess || hipEventCreate(&tl->e1) != hipSuccess)
    return d;  // untimed: this call runs the default plan
  t.next++;
  t.pending.push_back(tl);
  *timed = tl;
  return t.cand[tl->cand];
}

int tune_report(char *buf, size_t len) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  static const char *ops[] = {"conv",     "grad1",     "conv_grad1", "varlin",
                              "graddim",  "graddiff",  "shoot_vjp",  "hessian",
                              "varifold", "varifold_conv"};
  constexpr int nops = (int)(sizeof(ops) / sizeof(ops[0]));
  std::string out;
  for (auto &kv : g_tune) {
    resolve_pending(kv.second);
    const TuneKey &k = kv.first;
    const TuneEntry &t = kv.second;
    char line[256];
    const Choice &w = t.cand[t.done ? t.winner : 0];
    snprintf(line, sizeof(line),
             "op=%s prec=%d d=%d dv=%d batch=%d nx~2^%d ny~2^%d grid=%d -> R=%d waves=%d "
             "min_tiles=%d mx=%d w=%d %s %.4f ms (default %.4f ms) |",
             k.op >= 0 && k.op < nops ? ops[k.op] : "?", k.prec, k.D, k.DV, k.batch, k.lx, k.ly,
             k.grid, w.R, w.waves, w.min_tiles, w.mx, w.w,
             t.done ? "tuned" : "exploring", t.done ? t.best_ms[t.winner] : 0.0,
             t.best_ms.empty() ? 0.0 : t.best_ms[0]);
    out += line;
    for (size_t i = 0; i < t.cand.size(); i++) {
      const double ms = t.best_ms[i] < 3e38f ? t.best_ms[i] : -1.0;
      if (t.cand[i].w > 1)
        snprintf(line, sizeof(line), " %d/%d/%d/%d/w%d:%.3f", t.cand[i].R, t.cand[i].waves,
                 t.cand[i].min_tiles, t.cand[i].mx, t.cand[i].w, ms);
      else
        snprintf(line, sizeof(line), " %d/%d/%d/%d:%.3f", t.cand[i].R, t.cand[i].waves,
                 t.cand[i].min_tiles, t.cand[i].mx, ms);
      out += line;
    }
    out += "\n";
  }
  if (buf && len > 0) {
    const size_t n = std::min(len - 1, out.size());
    std::memcpy(buf, out.data(), n);
    buf[n] = 0;
  }
  return (int)out.size();
}

// Sort targets along a Hilbert curve; writes perm into the workspace.
int sort_targets(hipStream_t st, const double *x, long nx, int D, char *work, const Plan &pl) {
  double *part = (double *)(work + pl.off_bbox);
  uint32_t *kin = (uint32_t *)(work + pl.off_keys_in);
  uint32_t *kout = (uint32_t *)(work + pl.off_keys_out);
  uint32_t *vin = (uint32_t *)(work + pl.off_vals_in);
  uint32_t *perm = (uint32_t *)(work + pl.off_perm);
  const int nparts = (int)std::min<long>(kBBoxParts, std::max(1L, (nx + 255) / 256));
  hipLaunchKernelGGL(k_bbox_partial, dim3(nparts), dim3(256), 0, st, x, nx, D, part);
  hipLaunchKernelGGL(k_curve_keys, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, st, x, nx,
                     D, part, nparts, kin, vin);
  size_t tmp = pl.sort_tmp_bytes;
  hipError_t e = rocprim::radix_sort_pairs((void *)(work + pl.off_sort_tmp), tmp, kin, kout, vin,
                                           perm, (unsigned)nx, 0u, 30u, st);
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "radix sort: %s", hipGetErrorString(e));
  return DFM_OK;
}

int copy_d2d(hipStream_t st, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return DFM_OK;
  const bool aligned = ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0);
  const long n16 = aligned ? (long)(bytes / 16) : 0;
  const size_t done = (size_t)n16 * 16;
  const size_t rest = bytes - done;
  if (rest > 256) {  // unaligned buffers: let the runtime copy
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
    return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "memcpy: %s", hipGetErrorString(e));
  }
  const long blocks = std::max(1L, std::min<long>(8192, (n16 + 255) / 256));
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, st, (uint4 *)dst,
                     (const uint4 *)src, n16, (unsigned char *)dst + done,
                     (const unsigned char *)src + done, (int)rest);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "copy: %s", hipGetErrorString(e));
}

int fill(hipStream_t st, double *dst, long n, double v) {
  if (n <= 0 || !dst) return DFM_OK;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, n, v);
  return DFM_OK;
}

// The outputs of a call whose sums are all empty: zero, or the linear epilogue of zero.
static int store_empty(hipStream_t st, const KParams &p, long out0_n, long out1_n) {
  if (p.o_on) {
    const long n = std::max(out0_n, out1_n);
    if (n > 0)
      hipLaunchKernelGGL(k_put_empty, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p,
                         out0_n, out1_n);
  } else {
    fill(st, p.out0, out0_n, 0.0);
    fill(st, p.out1, out1_n, 0.0);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "fill: %s", hipGetErrorString(e));
}

int put_longs(hipStream_t st, long *dst, const long *src, int n) {
  for (int i = 0; i < n; i += 32) {
    LongChunk c{};
    const int k = std::min(32, n - i);
    for (int j = 0; j < k; j++) c.v[j] = src[i + j];
    hipLaunchKernelGGL(k_put_longs, dim3(1), dim3(64), 0, st, dst + i, c, k);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "offsets: %s", hipGetErrorString(e));
}

int seg_from_host(Seg &s, int n, const long *h, const long *d) {
  if (n < 1 || n > 65535) return set_error(DFM_ERR_ARG, "segment count must be in [1, 65535]");
  if (!h) return set_error(DFM_ERR_ARG, "segment offsets are NULL");
  if (h[0] != 0) return set_error(DFM_ERR_ARG, "segment offsets must start at 0");
  long mx = 0;
  for (int i = 0; i < n; i++) {
    const long c = h[i + 1] - h[i];
    if (c < 0) return set_error(DFM_ERR_ARG, "segment offsets must be non-decreasing");
    mx = std::max(mx, c);
  }
  if (h[n] > 0x7fffffffL) return set_error(DFM_ERR_ARG, "point counts must be in [0, 2^31)");
  s.n = n;
  s.h = h;
  s.d = d;
  s.total = h[n];
  s.max = mx;
  return DFM_OK;
}

// A fused integrator step finishes its source splits inside k_pairsum (a completion counter
// per target tile, whose arrivals each fence the device, writing back the L2) only on grids
// of fewer target tiles than this: the 1,000-point shoot's step (16 tiles) gains the launches
// it saves; the 4,096-point registration's step (64 tiles, 16 splits) took 0.062 ms in one
// launch against 0.037 ms for the matrix-core form with separate reduction and update
// launches, and the 64-subject atlas's batched step (2,048 tiles) 0.72 against 0.16 ms.
constexpr long kInKernelSplitTiles = 64;

// `pairs`: the call's (target, source) pairs, added to the profile's work with the timed
// launch.
static int launch_timed(hipStream_t st, int op, int prec, int D, int DV, KParams p,
                        const Plan &pl, char *work, TimedLaunch *timed, double pairs) {
  const double sigma = p.sigma;
  // FP64: K = exp(-|x' - y'|^2) with x' = x / sigma; FP32: K = exp2(-|x' - y'|^2) with
  // x' = x sqrt(log2 e) / sigma, so one v_exp_f32 evaluates the kernel.
  p.s = prec == DFM_FP64 ? 1.0 / sigma : std::sqrt(1.4426950408889634074) / sigma;
  p.exact_extent = kExactExtent;
  p.exact_extent_bf = pl.nseg > 1 ? kBfExtentBatch : kBfExtent;
  const double s2 = sigma * sigma;
  p.cg = -2.0 / (s2 * p.s);
  p.ch1 = 2.0 / s2;
  p.ch2 = -4.0 / (s2 * s2 * p.s * p.s);
  p.nsplit = pl.nsplit;
  p.split_len = pl.split_len;
  p.perm = nullptr;
  p.partial = pl.nsplit > 1 ? (double *)(work + pl.off_partial) : nullptr;
  if (pl.sort) {
    int rc = sort_targets(st, p.x, p.nx, D, work, pl);
    if (rc != DFM_OK) {
      if (timed) timed->state = 2;
      return rc;
    }
    p.perm = (const uint32_t *)(work + pl.off_perm);
  }
  // the tuner's events (created by tune_choose) or, when profiling, a pair of our own
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (timed) {
    ev0 = timed->e0;
    ev1 = timed->e1;
  } else if (g_profile.on) {
    if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess)
      return set_error(DFM_ERR_HIP, "hipEventCreate failed");
  }
  // k_pairsum finishes its own source splits and a fused integrator step where that is free
  // (one split) or cheap (grids of fewer than kInKernelSplitTiles target tiles); otherwise,
  // and always for the matrix-core kernels, k_reduce_splits adds the splits and
  // k_step_update applies the step after them
  const bool matrix_cores = prec == DFM_BF16MFMA ||
                            (prec == DFM_FP32C && pl.mx && mx_supported(op, D, DV) && pl.R == 1);
  const bool in_kernel =
      !matrix_cores && (pl.nsplit == 1 || pl.chunks * pl.nseg < kInKernelSplitTiles);
  const bool step_after = p.e_on && !in_kernel;
  if (step_after) p.e_on = 0;
  if (!in_kernel) p.done = nullptr;
  if (ev0) (void)hipEventRecord(ev0, st);
  int rc = prec == DFM_FP64       ? launch_op_f64(st, op, D, DV, p, pl)
           : prec == DFM_FP32X    ? launch_op_f32x(st, op, D, DV, p, pl)
           : prec == DFM_BF16MFMA ? launch_op_bf16(st, op, D, DV, p, pl)
                                  : launch_op_f32c(st, op, D, DV, p, pl);
  if (rc == DFM_OK && step_after) {
    const long rows = p.xoff ? pl.chunks * 64L * pl.R : p.nx;  // the largest segment, or all
    const dim3 g((unsigned)((rows + 255) / 256), (unsigned)pl.nseg);
    if (D == 2)
      hipLaunchKernelGGL(k_step_update<2>, g, dim3(256), 0, st, p);
    else
      hipLaunchKernelGGL(k_step_update<3>, g, dim3(256), 0, st, p);
  }
  hipError_t e = rc == DFM_OK ? hipGetLastError() : hipSuccess;
  if (ev1) (void)hipEventRecord(ev1, st);
  if (timed) {
    timed->state = (rc == DFM_OK && e == hipSuccess) ? 1 : 2;
  } else if (ev0) {
    g_profile.events.push_back({ev0, ev1});
    g_profile.pairs += pairs;
    g_profile.flops += pairs * flops_per_pair(op, D, DV);
  }
  if (rc != DFM_OK) return rc;
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "kernel launch: %s", hipGetErrorString(e));
  return DFM_OK;
}

// AbstractKernel::SetKernelWidth clamps the width at 1e-10 (AbstractKernel.h:85-91).
static double clamp_width(double h) { return h < 1e-10 ? 1e-10 : h; }

size_t pairsum_workspace(int op, int prec, int D, int DV, long nx, long ny) {
  if (nx <= 0 || ny <= 0) return 0;
  return make_plan(op, prec, D, DV, nx, ny).work_bytes;
}

size_t pairsum_batch_workspace(int op, int prec, int D, int DV, const Seg &xs, const Seg &ys) {
  if (xs.total <= 0 || ys.max <= 0) return 0;
  return make_batch_plan(op, prec, D, DV, xs.n, xs.total, xs.max, ys.max).work_bytes;
}

int pairsum_batch(hipStream_t st, int op, int prec, int D, int DV, double sigma, KParams p,
                  const Seg &xs, const Seg &ys, long out0_n, long out1_n, void *work,
                  size_t work_bytes) {
  if (xs.n != ys.n) return set_error(DFM_ERR_ARG, "target and source segment counts differ");
  prec = effective_prec(op, prec, D, DV, xs.n, xs.max, ys.max);
  p.nx = xs.total;
  p.ny = ys.total;
  p.xoff = xs.n > 1 ? xs.d : nullptr;
  p.yoff = ys.n > 1 ? ys.d : nullptr;
  if (xs.n > 1 && (!xs.d || !ys.d)) return set_error(DFM_ERR_ARG, "device segment offsets missing");
  if (p.nx == 0) return DFM_OK;
  if (ys.max == 0) return store_empty(st, p, out0_n, out1_n);  // every source set is empty
  std::shared_ptr<TimedLaunch> timed;
  const Choice c = tune_choose(op, prec, D, DV, 1, xs.n, xs.max, ys.max, &timed);
  Plan pl = batch_plan(op, prec, D, DV, xs.n, xs.total, xs.max, ys.max, c);
  if (pl.work_bytes > 0 && (!work || work_bytes < pl.work_bytes)) {
    if (timed) timed->state = 2;
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes, got %zu", pl.work_bytes,
                     work ? work_bytes : (size_t)0);
  }
  p.sigma = clamp_width(sigma);
  double pairs = 0.0;
  for (int i = 0; i < xs.n; i++) pairs += (double)xs.count(i) * (double)ys.count(i);
  return launch_timed(st, op, prec, D, DV, p, pl, (char *)work, timed.get(), pairs);
}

int pairsum(hipStream_t st, int op, int prec, int D, int DV, double sigma, KParams p,
            long out0_n, long out1_n, void *work, size_t work_bytes, const char *ws_label) {
  if (p.nx == 0) return DFM_OK;
  prec = effective_prec(op, prec, D, DV, 1, p.nx, p.ny);
  if (p.ny == 0) return store_empty(st, p, out0_n, out1_n);  // empty sums
  std::shared_ptr<TimedLaunch> timed;
  const Choice c = tune_choose(op, prec, D, DV, 0, 1, p.nx, p.ny, &timed);
  Plan pl = single_plan(op, prec, D, DV, p.nx, p.ny, c);
  if (pl.work_bytes > 0 && (!work || work_bytes < pl.work_bytes)) {
    if (timed) timed->state = 2;
    return set_error(DFM_ERR_ARG, "%s too small: need %zu bytes, got %zu", ws_label,
                     pl.work_bytes, work ? work_bytes : (size_t)0);
  }
  p.sigma = clamp_width(sigma);
  return launch_timed(st, op, prec, D, DV, p, pl, (char *)work, timed.get(),
                      (double)p.nx * (double)p.ny);
}

}  // namespace dfm
