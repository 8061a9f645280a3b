// pairsum.hip — host side of the pair-sum engine: per-thread error and timing state,
// the Morton ordering of the targets (bounding box, 30-bit keys, rocPRIM radix sort),
// the work plan and workspace layout of a call, and the launch of the per-precision
// kernel instantiations (pairsum_f64.hip, pairsum_f32c.hip, pairsum_f32x.hip).
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

// Device-to-device copy of 16-byte groups (plus a byte tail); also the known-size
// reference kernel for scaling HBM traffic counters.
__global__ __launch_bounds__(256) void k_copy(uint4 *dst, const uint4 *src, long n16,
                                              unsigned char *dtail, const unsigned char *stail,
                                              int ntail) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) dtail[threadIdx.x] = stail[threadIdx.x];
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

// ------------------------------------------------------------------------------------
// planning

static int targets_per_lane(int op, int prec, int D, int DV) {
  if (prec == DFM_FP64) return 1;
  switch (op) {
    case OP_VARLIN: return DV * D > 9 ? 1 : 2;
    case OP_GRADDIFF: return 2;
    case OP_SHOOTVJP: return 1;
    default: return 4;
  }
}

static int tile_of(int prec) { return prec == DFM_FP64 ? kTileF64 : kTileF32; }

static int nacc_of(int op, int D, int DV) {
  switch (op) {
    case OP_CONV: return DV;
    case OP_GRAD1: return D;
    case OP_CONVGRAD1: return DV + D;
    case OP_VARLIN: return DV * D;
    case OP_GRADDIM: return DV;
    case OP_GRADDIFF: return 2 * D;
    default: return 5 * D;
  }
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Work split shared by single calls and ragged batches: chunks of the largest segment's
// targets, and source splits of the largest segment's sources chosen so that the grid
// holds >= 4096 wavefronts (4 per SIMD on 256 CUs) with >= 4 source tiles per split.
static void split_plan(Plan &pl, int op, int prec, int D, int DV, int nseg, long max_nx,
                       long max_ny) {
  pl.R = targets_per_lane(op, prec, D, DV);
  pl.nseg = nseg;
  pl.chunks = (max_nx + 64L * pl.R - 1) / (64L * pl.R);
  const int TS = tile_of(prec);
  const long tiles = (max_ny + TS - 1) / TS;
  const long blocks = pl.chunks * nseg;
  long want = blocks > 0 ? (4096 + blocks - 1) / blocks : 1;
  long maxsplit = std::max(1L, tiles / 4);
  long ns = std::max(1L, std::min(want, maxsplit));
  ns = std::min(ns, 65535L);
  const long per = (tiles + ns - 1) / ns;
  pl.split_len = std::max(1L, per) * TS;
  pl.nsplit = (int)std::max(1L, (max_ny + pl.split_len - 1) / pl.split_len);
}

Plan make_batch_plan(int op, int prec, int D, int DV, int nseg, long nx_total, long max_nx,
                     long max_ny) {
  Plan pl{};
  split_plan(pl, op, prec, D, DV, nseg, max_nx, max_ny);
  pl.sort = false;
  size_t off = 0;
  if (pl.nsplit > 1) {
    pl.off_partial = off;
    off = align256(off + sizeof(double) * (size_t)pl.nsplit * nx_total * nacc_of(op, D, DV));
  }
  pl.work_bytes = off;
  return pl;
}

Plan make_plan(int op, int prec, int D, int DV, long nx, long ny) {
  Plan pl{};
  split_plan(pl, op, prec, D, DV, 1, nx, ny);
  pl.sort = prec != DFM_FP64;
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

int launch_pairsum(hipStream_t st, int op, int prec, int D, int DV, KParams p,
                   const Plan &pl, char *work) {
  const double sigma = p.sigma;
  // FP64: K = exp(-|x' - y'|^2) with x' = x / sigma; FP32: K = exp2(-|x' - y'|^2) with
  // x' = x sqrt(log2 e) / sigma, so one v_exp_f32 evaluates the kernel.
  p.s = prec == DFM_FP64 ? 1.0 / sigma : std::sqrt(1.4426950408889634074) / sigma;
  p.exact_extent = kExactExtent;
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
    if (rc != DFM_OK) return rc;
    p.perm = (const uint32_t *)(work + pl.off_perm);
  }
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (g_profile.on) {
    if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess)
      return set_error(DFM_ERR_HIP, "hipEventCreate failed");
    (void)hipEventRecord(ev0, st);
  }
  int rc = prec == DFM_FP64    ? launch_op_f64(st, op, D, DV, p, pl)
           : prec == DFM_FP32X ? launch_op_f32x(st, op, D, DV, p, pl)
                               : launch_op_f32c(st, op, D, DV, p, pl);
  if (g_profile.on) {
    (void)hipEventRecord(ev1, st);
    g_profile.events.push_back({ev0, ev1});
  }
  if (rc != DFM_OK) return rc;
  hipError_t e = hipGetLastError();
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
  p.nx = xs.total;
  p.ny = ys.total;
  p.xoff = xs.n > 1 ? xs.d : nullptr;
  p.yoff = ys.n > 1 ? ys.d : nullptr;
  if (xs.n > 1 && (!xs.d || !ys.d)) return set_error(DFM_ERR_ARG, "device segment offsets missing");
  if (p.nx == 0) return DFM_OK;
  if (ys.max == 0) {  // every segment has an empty source set
    fill(st, p.out0, out0_n, 0.0);
    fill(st, p.out1, out1_n, 0.0);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "fill: %s", hipGetErrorString(e));
  }
  Plan pl = make_batch_plan(op, prec, D, DV, xs.n, xs.total, xs.max, ys.max);
  if (pl.work_bytes > 0 && (!work || work_bytes < pl.work_bytes))
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes, got %zu", pl.work_bytes,
                     work ? work_bytes : (size_t)0);
  p.sigma = clamp_width(sigma);
  return launch_pairsum(st, op, prec, D, DV, p, pl, (char *)work);
}

int pairsum(hipStream_t st, int op, int prec, int D, int DV, double sigma, KParams p,
            long out0_n, long out1_n, void *work, size_t work_bytes) {
  if (p.nx == 0) return DFM_OK;
  if (p.ny == 0) {  // empty sums
    fill(st, p.out0, out0_n, 0.0);
    fill(st, p.out1, out1_n, 0.0);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "fill: %s", hipGetErrorString(e));
  }
  Plan pl = make_plan(op, prec, D, DV, p.nx, p.ny);
  if (pl.work_bytes > 0 && (!work || work_bytes < pl.work_bytes))
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes, got %zu", pl.work_bytes,
                     work ? work_bytes : (size_t)0);
  p.sigma = clamp_width(sigma);
  return launch_pairsum(st, op, prec, D, DV, p, pl, (char *)work);
}

}  // namespace dfm
