// dfm_internal.h — types shared by the kernel file and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace dfm {

enum : int {
  OP_CONV = 0,
  OP_GRAD1 = 1,
  OP_CONVGRAD1 = 2,
  OP_VARLIN = 3,
  OP_GRADDIM = 4,
  OP_GRADDIFF = 5,
  OP_SHOOTVJP = 6
};

constexpr int kBBoxParts = 256;

// Source tile (LDS) per wavefront and, for FP32, sources summed in FP32 before each fold
// into the FP64 accumulators.
constexpr int kTileF64 = 128;
constexpr int kTileF32 = 256;
constexpr int kFoldF32 = 32;

// FP32 paths: a wavefront whose targets span more than this many scaled kernel units
// (sqrt(log2 e) / sigma per unit length) forms exact (hi + lo) coordinate differences.
constexpr double kExactExtent = 4.0;

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
};

// Work decomposition and workspace layout of one call.
struct Plan {
  int R;            // targets per lane
  long chunks;      // workgroups along the targets
  int nsplit;       // source splits
  long split_len;   // sources per split (multiple of the tile)
  size_t off_bbox, off_keys_in, off_keys_out, off_vals_in, off_perm, off_sort_tmp, off_partial;
  size_t sort_tmp_bytes;
  size_t work_bytes;
};

int set_error(int code, const char *fmt, ...);
const char *last_error();
void clear_error();

Plan make_plan(int op, int prec, int D, int DV, long nx, long ny);
int launch_pairsum(hipStream_t st, int op, int prec, int D, int DV, KParams p, const Plan &pl,
                   char *work);
int fill(hipStream_t st, double *dst, long n, double v);

// One pair-sum call on device pointers: plans the work, checks the workspace, clamps the
// kernel width and launches (empty source sets give zero outputs).
int pairsum(hipStream_t st, int op, int prec, int D, int DV, double sigma, KParams p,
            long out0_n, long out1_n, void *work, size_t work_bytes);
size_t pairsum_workspace(int op, int prec, int D, int DV, long nx, long ny);
int copy_d2d(hipStream_t st, void *dst, const void *src, size_t bytes);

// one per precision (pairsum_f64.hip, pairsum_f32c.hip, pairsum_f32x.hip)
int launch_op_f64(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl);
int launch_op_f32c(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl);
int launch_op_f32x(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl);
void profile_enable(bool on);
int profile_read(double *total_ms, long *launches);

}  // namespace dfm
