// pairsum_f32c.hip — instantiates the pair-sum kernels for FP32 with per-wavefront exact
// differences where needed, and the matrix-core form of the fused Convolve + Grad1.
#include "pairsum_mfma.h"

namespace dfm {
namespace {

// Compact wavefronts on the matrix cores, wide ones on k_pairsum (R = 1, exact path), then
// the split reduction: every target is written by exactly one of the two kernels.
template <int OP, int DV>
void launch_mx(hipStream_t st, const KParams &p, const Plan &pl) {
  dim3 grid((unsigned)pl.chunks, (unsigned)pl.nsplit, (unsigned)pl.nseg);
  hipLaunchKernelGGL((k_pairsum_mx<OP, DV>), grid, dim3(64), 0, st, p);
  KParams q = p;
  q.wave_filter = 1;
  hipLaunchKernelGGL((k_pairsum<float, OP, 3, DV, 0, 1, MODE_ADAPTIVE, 1>), grid, dim3(64), 0, st, q);
  if (pl.nsplit > 1)
    hipLaunchKernelGGL((k_reduce_splits<OP, 3, DV>), dim3((unsigned)((p.nx + 255) / 256)),
                       dim3(256), 0, st, p);
}

}  // namespace

bool mx_supported(int op, int D, int DV) {
  return D == 3 && (op == OP_CONVGRAD1 || op == OP_GRAD1) && (DV == 3 || DV == 6);
}

int launch_op_f32c(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl) {
  if (pl.mx && mx_supported(op, D, DV) && pl.R == 1) {
    if (op == OP_CONVGRAD1) {
      if (DV == 3) launch_mx<OP_CONVGRAD1, 3>(st, p, pl);
      else launch_mx<OP_CONVGRAD1, 6>(st, p, pl);
    } else {
      if (DV == 3) launch_mx<OP_GRAD1, 3>(st, p, pl);
      else launch_mx<OP_GRAD1, 6>(st, p, pl);
    }
    return DFM_OK;
  }
  return launch_op<float, MODE_ADAPTIVE>(st, op, D, DV, p, pl);
}

}  // namespace dfm
