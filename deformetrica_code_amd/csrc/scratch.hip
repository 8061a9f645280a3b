// scratch.hip — dfm_conv_grad1_noalloc: the fused Convolve + ConvolveGradient(X, alpha) of
// dfm_conv_grad1 with its workspace taken from a file-scope device array instead of a
// caller buffer, for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` has one copy per GPU, resident until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in, not a device
// address): the first call on a device, with that device current, looks up the device
// address with hipGetSymbolAddress and keeps it per device; every launch then uses that
// address.  The call runs the same plan as dfm_conv_grad1 at the same sizes (heuristic or
// tuned), so its results are bitwise those of dfm_conv_grad1.  Nothing in the scratch must
// start at zero: the Hilbert keys, permutation and split partials a call reads are all
// written earlier in the same call.  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "dfm_hip.h"
#include "dfm_internal.h"

using namespace dfm;

static __device__ __align__(256) char scratch[DFM_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch` on each device, once looked up

// Device address of the current device's copy of `scratch` (looked up on first use).
int scratch_address(void **out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(DFM_ERR_ARG, "device %d is beyond the %d devices the static scratch tracks",
                     dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(DFM_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DFM_OK;
}

}  // namespace

extern "C" int dfm_conv_grad1_noalloc(void *stream, int D, int DV, double sigma, int prec,
                                      const double *x, const double *alpha, long nx,
                                      const double *y, const double *beta, long ny,
                                      double *conv, double *grad) {
  clear_error();
  // the argument checks of dfm_conv_grad1
  if (D != 2 && D != 3) return set_error(DFM_ERR_DIM, "unsupported dimpoint %d (2 or 3)", D);
  if (DV < 1 || DV > 6) return set_error(DFM_ERR_DIM, "unsupported dimvect %d (1..6)", DV);
  if (!valid_precision(prec))
    return set_error(DFM_ERR_ARG, "unknown precision %d", prec);
  if (!(sigma == sigma)) return set_error(DFM_ERR_ARG, "kernel width is NaN");
  if (nx < 0 || ny < 0 || nx > 0x7fffffffL || ny > 0x7fffffffL)
    return set_error(DFM_ERR_ARG, "point counts must be in [0, 2^31)");
  if (nx > 0 && !x) return set_error(DFM_ERR_ARG, "x is NULL");
  if (nx > 0 && grad && !alpha) return set_error(DFM_ERR_ARG, "alpha is NULL");
  if (nx > 0 && ny > 0 && (!y || !beta)) return set_error(DFM_ERR_ARG, "y or beta is NULL");
  if (!conv && !grad) return DFM_OK;

  void *work = nullptr;
  int rc = scratch_address(&work);
  if (rc != DFM_OK) return rc;
  const char *label = "static scratch (DFM_STATIC_SCRATCH_BYTES)";
  hipStream_t st = (hipStream_t)stream;
  KParams p;
  memset(&p, 0, sizeof(p));
  p.x = x; p.ta = alpha; p.y = y; p.w = beta; p.nx = nx; p.ny = ny;
  if (!grad) {
    p.out0 = conv;
    return pairsum(st, OP_CONV, prec, D, DV, sigma, p, nx * DV, 0, work, DFM_STATIC_SCRATCH_BYTES,
                   label);
  }
  if (!conv) {
    p.out0 = grad;
    return pairsum(st, OP_GRAD1, prec, D, DV, sigma, p, nx * D, 0, work, DFM_STATIC_SCRATCH_BYTES,
                   label);
  }
  p.out0 = conv;
  p.out1 = grad;
  return pairsum(st, OP_CONVGRAD1, prec, D, DV, sigma, p, nx * DV, nx * D, work,
                 DFM_STATIC_SCRATCH_BYTES, label);
}
