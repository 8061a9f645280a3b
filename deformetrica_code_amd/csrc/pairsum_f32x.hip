// pairsum_f32x.hip — instantiates the pair-sum kernels for FP32 with exact differences everywhere.
#include "pairsum_kernels.h"

namespace dfm {

int launch_op_f32x(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl) {
  return launch_op<float, MODE_EXACT>(st, op, D, DV, p, pl);
}

}  // namespace dfm
