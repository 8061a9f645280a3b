// pairsum_f32c.hip — instantiates the pair-sum kernels for FP32 with per-wavefront exact differences where needed.
#include "pairsum_kernels.h"

namespace dfm {

int launch_op_f32c(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl) {
  return launch_op<float, MODE_ADAPTIVE>(st, op, D, DV, p, pl);
}

}  // namespace dfm
