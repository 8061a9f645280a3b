// pairsum_bf16.hip — launches of the DFM_BF16MFMA pair sums: the matrix-core kernel
// (k_pairsum_bf; its wide groups of 256 targets take the FP32C wavefront path in the same
// launch), then the split reduction.  Single calls (Hilbert-sorted targets) and ragged
// batches (grid.z = segment, the caller's target order) alike.
#include "pairsum_bf16.h"

namespace dfm {
namespace {

template <int OP, int D, int DV>
void launch_bf(hipStream_t st, const KParams &p, const Plan &pl) {
  const dim3 gbf((unsigned)pl.chunks, (unsigned)pl.nsplit, (unsigned)pl.nseg);
  hipLaunchKernelGGL((k_pairsum_bf<OP, D, DV, kBfGroup / 64>), gbf, dim3(256), 0, st, p);
  if (pl.nsplit > 1)
    hipLaunchKernelGGL((k_reduce_splits<OP, D, DV>), dim3((unsigned)((p.nx + 255) / 256)),
                       dim3(256), 0, st, p);
}

template <int OP, int D>
int launch_bf_dv(hipStream_t st, int DV, const KParams &p, const Plan &pl) {
  switch (DV) {
    case 1: launch_bf<OP, D, 1>(st, p, pl); return DFM_OK;
    case 2: launch_bf<OP, D, 2>(st, p, pl); return DFM_OK;
    case 3: launch_bf<OP, D, 3>(st, p, pl); return DFM_OK;
    case 4: launch_bf<OP, D, 4>(st, p, pl); return DFM_OK;
    case 5: launch_bf<OP, D, 5>(st, p, pl); return DFM_OK;
    case 6: launch_bf<OP, D, 6>(st, p, pl); return DFM_OK;
  }
  return set_error(DFM_ERR_DIM, "unsupported dimvect %d", DV);
}

template <int OP>
int launch_bf_d(hipStream_t st, int D, int DV, const KParams &p, const Plan &pl) {
  if constexpr (OP == OP_VARIFOLD || OP == OP_VARIFOLD_CONV) {  // weights u of dimension D
    if (DV != D) return set_error(DFM_ERR_DIM, "the varifold sums need dimvect == dimpoint");
    if (D == 2) {
      launch_bf<OP, 2, 2>(st, p, pl);
      return DFM_OK;
    }
    if (D == 3) {
      launch_bf<OP, 3, 3>(st, p, pl);
      return DFM_OK;
    }
  } else {
    if (D == 2) return launch_bf_dv<OP, 2>(st, DV, p, pl);
    if (D == 3) return launch_bf_dv<OP, 3>(st, DV, p, pl);
  }
  return set_error(DFM_ERR_DIM, "unsupported dimpoint %d", D);
}

}  // namespace

bool bf_supported(int op, int D, int DV) {
  if (op == OP_VARIFOLD || op == OP_VARIFOLD_CONV) return (D == 2 || D == 3) && DV == D;
  return (op == OP_CONV || op == OP_GRAD1 || op == OP_CONVGRAD1) && (D == 2 || D == 3) &&
         DV >= 1 && DV <= 6;
}

int launch_op_bf16(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl) {
  if (!bf_supported(op, D, DV) || pl.R != kBfGroup / 64)
    return set_error(DFM_ERR_ARG, "the BF16 cross-term kernel takes Convolve / Grad1 / fused / "
                                  "varifold sums on plans of %d-target groups", kBfGroup);
  switch (op) {
    case OP_CONV: return launch_bf_d<OP_CONV>(st, D, DV, p, pl);
    case OP_GRAD1: return launch_bf_d<OP_GRAD1>(st, D, DV, p, pl);
    case OP_VARIFOLD: return launch_bf_d<OP_VARIFOLD>(st, D, DV, p, pl);
    case OP_VARIFOLD_CONV: return launch_bf_d<OP_VARIFOLD_CONV>(st, D, DV, p, pl);
    default: return launch_bf_d<OP_CONVGRAD1>(st, D, DV, p, pl);
  }
}

}  // namespace dfm
