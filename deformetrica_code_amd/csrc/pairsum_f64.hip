// pairsum_f64.hip — instantiates the pair-sum kernels for FP64 (the reference arithmetic).
#include "pairsum_kernels.h"

namespace dfm {

int launch_op_f64(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl) {
  return launch_op<double, MODE_ADAPTIVE>(st, op, D, DV, p, pl);
}

}  // namespace dfm
