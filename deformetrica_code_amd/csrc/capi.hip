// capi.hip — the extern "C" boundary of the kernel-sum engine (declared in dfm_hip.h).
//
// Device entry points validate their arguments the way the reference plugin does
// (CUDAExactKernel.cxx: row-count mismatch, unsupported dimvect, dim out of range) and
// return a status code; host-pointer wrappers take the place of the reference's
// GaussGpu*Conv1D host functions (GpuConv1D.cu:88-133 and siblings), which allocate, copy
// in, launch, synchronise, copy out and free on every call: here the device buffers are
// leased from a pool that keeps them resident between calls (grown when a call needs more,
// never freed per call), so a call is copy in, launch, copy out and one synchronisation of
// the calling thread's stream — with every HIP error checked.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <vector>

#include "dfm_hip.h"
#include "dfm_internal.h"

using namespace dfm;

namespace {

constexpr long kMaxCount = 0x7fffffffL;

int check_dims(int D, int DV) {
  if (D != 2 && D != 3) return set_error(DFM_ERR_DIM, "unsupported dimpoint %d (2 or 3)", D);
  if (DV < 1 || DV > 6) return set_error(DFM_ERR_DIM, "unsupported dimvect %d (1..6)", DV);
  return DFM_OK;
}

int check_common(int D, int DV, int prec, double sigma, long nx, long ny) {
  int rc = check_dims(D, DV);
  if (rc) return rc;
  if (!valid_precision(prec))
    return set_error(DFM_ERR_ARG, "unknown precision %d", prec);
  if (!(sigma == sigma)) return set_error(DFM_ERR_ARG, "kernel width is NaN");
  if (nx < 0 || ny < 0 || nx > kMaxCount || ny > kMaxCount)
    return set_error(DFM_ERR_ARG, "point counts must be in [0, 2^31)");
  return DFM_OK;
}

int need(const void *ptr, const char *name) {
  if (!ptr) return set_error(DFM_ERR_ARG, "%s is NULL", name);
  return DFM_OK;
}

// Common body of every device entry point.
int run(void *stream, int op, int D, int DV, double sigma, int prec, KParams p, long out0_n,
        long out1_n, void *work, size_t work_bytes) {
  clear_error();
  return pairsum((hipStream_t)stream, op, prec, D, DV, sigma, p, out0_n, out1_n, work,
                 work_bytes);
}

KParams blank() {
  KParams p;
  std::memset(&p, 0, sizeof(p));
  return p;
}

}  // namespace

extern "C" {

int dfm_abi_version(void) { return DFM_ABI_VERSION; }

const char *dfm_last_error(void) { return last_error(); }

int dfm_device_count(int *count) {
  clear_error();
  if (!count) return set_error(DFM_ERR_ARG, "count is NULL");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return set_error(DFM_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  return DFM_OK;
}

int dfm_malloc(void **ptr, size_t bytes) {
  clear_error();
  if (!ptr) return set_error(DFM_ERR_ARG, "ptr is NULL");
  *ptr = nullptr;
  if (bytes == 0) return DFM_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return set_error(DFM_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return DFM_OK;
}

int dfm_free(void *ptr) {
  clear_error();
  if (!ptr) return DFM_OK;
  hipError_t e = hipFree(ptr);
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "hipFree: %s", hipGetErrorString(e));
}

static int memcpy_on(void *dst, const void *src, size_t bytes, int kind, void *stream, bool wait) {
  clear_error();
  if (bytes == 0) return DFM_OK;
  if (!dst || !src) return set_error(DFM_ERR_ARG, "NULL pointer");
  if (kind < 0 || kind > 2) return set_error(DFM_ERR_ARG, "unknown copy kind %d", kind);
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (kind == 2) {
    int rc = copy_d2d(st, dst, src, bytes);
    if (rc != DFM_OK) return rc;
    e = wait ? hipStreamSynchronize(st) : hipSuccess;
  } else {
    e = hipMemcpyAsync(dst, src, bytes, k, st);
    if (e == hipSuccess && wait) e = hipStreamSynchronize(st);
  }
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "memcpy: %s", hipGetErrorString(e));
}

int dfm_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream) {
  return memcpy_on(dst, src, bytes, kind, stream, true);
}

int dfm_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream) {
  return memcpy_on(dst, src, bytes, kind, stream, false);
}

int dfm_stream_create(void **stream) {
  clear_error();
  if (!stream) return set_error(DFM_ERR_ARG, "stream is NULL");
  hipStream_t st = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  *stream = (void *)st;
  return DFM_OK;
}

int dfm_stream_destroy(void *stream) {
  clear_error();
  if (!stream) return DFM_OK;
  hipError_t e = hipStreamDestroy((hipStream_t)stream);
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "hipStreamDestroy: %s", hipGetErrorString(e));
}

int dfm_stream_synchronize(void *stream) {
  clear_error();
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
}

int dfm_graph_begin_capture(void *stream) {
  clear_error();
  if (!stream) return set_error(DFM_ERR_ARG, "capture needs a created stream, not the default one");
  hipError_t e = hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
}

int dfm_graph_end_capture(void *stream, void **graph_exec) {
  clear_error();
  if (!stream || !graph_exec) return set_error(DFM_ERR_ARG, "NULL argument");
  *graph_exec = nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  *graph_exec = (void *)ex;
  return DFM_OK;
}

int dfm_graph_launch(void *graph_exec, void *stream) {
  clear_error();
  if (!graph_exec) return set_error(DFM_ERR_ARG, "graph is NULL");
  hipError_t e = hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream);
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(e));
}

int dfm_graph_destroy(void *graph_exec) {
  clear_error();
  if (!graph_exec) return DFM_OK;
  hipError_t e = hipGraphExecDestroy((hipGraphExec_t)graph_exec);
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "hipGraphExecDestroy: %s", hipGetErrorString(e));
}

int dfm_autotune(int mode) {
  clear_error();
  if (mode < 0 || mode > 2) return set_error(DFM_ERR_ARG, "autotune mode must be 0, 1 or 2");
  tune_set_mode(mode);
  return DFM_OK;
}

int dfm_autotune_reset(void) {
  clear_error();
  tune_reset();
  return DFM_OK;
}

int dfm_autotune_report(char *buf, size_t len) { return tune_report(buf, len); }

int dfm_profile_enable(int on) {
  profile_enable(on != 0);
  return DFM_OK;
}

int dfm_profile_read(double *total_ms, long *launches) {
  clear_error();
  return profile_read(total_ms, launches);
}

int dfm_profile_work(double *pairs, double *flops) {
  clear_error();
  return profile_work(pairs, flops);
}

double dfm_flops_per_pair(int op, int dimpoint, int dimvect) {
  return flops_per_pair(op, dimpoint, dimvect);
}

size_t dfm_workspace_size(int op, int dimpoint, int dimvect, long nx, long ny, int precision) {
  if (check_common(dimpoint, dimvect, precision, 1.0, nx, ny) != DFM_OK) return 0;
  if (op < DFM_OP_CONV || op > DFM_OP_VARIFOLD_CONV) return 0;
  if ((op == DFM_OP_VARIFOLD || op == DFM_OP_VARIFOLD_CONV) && dimvect != dimpoint) return 0;
  if (op == DFM_OP_GRADDIFF || op == DFM_OP_SHOOT_VJP) ny = nx;
  if (nx == 0 || ny == 0) return 0;
  return make_plan(op, precision, dimpoint, dimvect, nx, ny).work_bytes;
}

int dfm_conv(void *stream, int D, int DV, double sigma, int prec, const double *x, long nx,
             const double *y, const double *beta, long ny, double *gamma, void *work,
             size_t work_bytes) {
  int rc = check_common(D, DV, prec, sigma, nx, ny);
  if (!rc && nx > 0) rc = need(x, "x") ? DFM_ERR_ARG : need(gamma, "gamma");
  if (!rc && nx > 0 && ny > 0) rc = need(y, "y") ? DFM_ERR_ARG : need(beta, "beta");
  if (rc) return rc;
  KParams p = blank();
  p.x = x; p.y = y; p.w = beta; p.out0 = gamma; p.nx = nx; p.ny = ny;
  return run(stream, OP_CONV, D, DV, sigma, prec, p, nx * DV, 0, work, work_bytes);
}

int dfm_grad1(void *stream, int D, int DV, double sigma, int prec, const double *x,
              const double *alpha, long nx, const double *y, const double *beta, long ny,
              double *gamma, void *work, size_t work_bytes) {
  int rc = check_common(D, DV, prec, sigma, nx, ny);
  if (!rc && nx > 0)
    rc = (need(x, "x") || need(alpha, "alpha")) ? DFM_ERR_ARG : need(gamma, "gamma");
  if (!rc && nx > 0 && ny > 0) rc = need(y, "y") ? DFM_ERR_ARG : need(beta, "beta");
  if (rc) return rc;
  KParams p = blank();
  p.x = x; p.ta = alpha; p.y = y; p.w = beta; p.out0 = gamma; p.nx = nx; p.ny = ny;
  return run(stream, OP_GRAD1, D, DV, sigma, prec, p, nx * D, 0, work, work_bytes);
}

int dfm_conv_grad1(void *stream, int D, int DV, double sigma, int prec, const double *x,
                   const double *alpha, long nx, const double *y, const double *beta, long ny,
                   double *conv, double *grad, void *work, size_t work_bytes) {
  int rc = check_common(D, DV, prec, sigma, nx, ny);
  if (!rc && nx > 0) rc = need(x, "x");
  if (!rc && nx > 0 && grad) rc = need(alpha, "alpha");
  if (!rc && nx > 0 && ny > 0) rc = need(y, "y") ? DFM_ERR_ARG : need(beta, "beta");
  if (rc) return rc;
  if (!conv && !grad) return DFM_OK;
  KParams p = blank();
  p.x = x; p.ta = alpha; p.y = y; p.w = beta; p.out0 = conv; p.out1 = grad; p.nx = nx; p.ny = ny;
  if (!grad) return run(stream, OP_CONV, D, DV, sigma, prec, p, nx * DV, 0, work, work_bytes);
  if (!conv) {
    p.out0 = grad;
    p.out1 = nullptr;
    return run(stream, OP_GRAD1, D, DV, sigma, prec, p, nx * D, 0, work, work_bytes);
  }
  return run(stream, OP_CONVGRAD1, D, DV, sigma, prec, p, nx * DV, nx * D, work, work_bytes);
}

int dfm_varlin(void *stream, int D, int DV, double sigma, int prec, const double *x, long nx,
               const double *y, const double *beta, long ny, double *gamma, void *work,
               size_t work_bytes) {
  int rc = check_common(D, DV, prec, sigma, nx, ny);
  if (!rc && nx > 0) rc = need(x, "x") ? DFM_ERR_ARG : need(gamma, "gamma");
  if (!rc && nx > 0 && ny > 0) rc = need(y, "y") ? DFM_ERR_ARG : need(beta, "beta");
  if (rc) return rc;
  KParams p = blank();
  p.x = x; p.y = y; p.w = beta; p.out0 = gamma; p.nx = nx; p.ny = ny;
  return run(stream, OP_VARLIN, D, DV, sigma, prec, p, nx * DV * D, 0, work, work_bytes);
}

int dfm_hessian(void *stream, int D, int DV, double sigma, int prec, const double *x, long nx,
                const double *y, const double *beta, long ny, double *gamma, void *work,
                size_t work_bytes) {
  int rc = check_common(D, DV, prec, sigma, nx, ny);
  if (!rc && nx > 0) rc = need(x, "x") ? DFM_ERR_ARG : need(gamma, "gamma");
  if (!rc && nx > 0 && ny > 0) rc = need(y, "y") ? DFM_ERR_ARG : need(beta, "beta");
  if (rc) return rc;
  KParams p = blank();
  p.x = x; p.y = y; p.w = beta; p.out0 = gamma; p.nx = nx; p.ny = ny;
  return run(stream, OP_HESS, D, DV, sigma, prec, p, nx * DV * D * D, 0, work, work_bytes);
}

int dfm_graddim(void *stream, int D, int DV, double sigma, int prec, const double *x, long nx,
                const double *y, const double *beta, long ny, int dim, double *gamma, void *work,
                size_t work_bytes) {
  int rc = check_common(D, DV, prec, sigma, nx, ny);
  if (!rc && (dim < 0 || dim >= D))
    rc = set_error(DFM_ERR_INDEX, "dimension index out of bounds");
  if (!rc && nx > 0) rc = need(x, "x") ? DFM_ERR_ARG : need(gamma, "gamma");
  if (!rc && nx > 0 && ny > 0) rc = need(y, "y") ? DFM_ERR_ARG : need(beta, "beta");
  if (rc) return rc;
  KParams p = blank();
  p.x = x; p.y = y; p.w = beta; p.out0 = gamma; p.nx = nx; p.ny = ny; p.dim = dim;
  return run(stream, OP_GRADDIM, D, DV, sigma, prec, p, nx * DV, 0, work, work_bytes);
}

int dfm_graddiff(void *stream, int D, int DV, double sigma, int prec, const double *x,
                 const double *beta, const double *eta, long n, double *gamma, void *work,
                 size_t work_bytes) {
  int rc = check_common(D, DV, prec, sigma, n, n);
  if (!rc && DV != D)
    rc = set_error(DFM_ERR_DIM, "Invalid number of columns of beta (needs dimvect == dimpoint)");
  if (!rc && n > 0)
    rc = (need(x, "x") || need(beta, "beta") || need(eta, "eta")) ? DFM_ERR_ARG
                                                                  : need(gamma, "gamma");
  if (rc) return rc;
  KParams p = blank();
  p.x = x; p.ta = beta; p.tb = eta; p.y = x; p.w = beta; p.se = eta; p.out0 = gamma;
  p.nx = n; p.ny = n;
  return run(stream, OP_GRADDIFF, D, DV, sigma, prec, p, n * D, 0, work, work_bytes);
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// host-pointer wrappers

namespace {

// A set of device buffers and one pinned host staging buffer for one host-pointer call.
// Arenas live in a per-device pool for the life of the process: a call leases one (another
// thread's concurrent call leases another), and a buffer is reallocated only when a call
// needs more than it holds, so a steady sequence of calls (a shoot, a subject loop)
// allocates and frees nothing.  A call packs its inputs into the staging buffer for one
// host-to-device copy, and its outputs come back in one device-to-host copy.
struct Arena {
  static constexpr int kSlots = 4;  // packed inputs, packed outputs, workspace
  int device = 0;
  void *p[kSlots] = {};
  size_t cap[kSlots] = {};
  void *host = nullptr;      // pinned and mapped (hipHostMalloc)
  void *host_dev = nullptr;  // its device address
  size_t host_cap = 0;
};


std::mutex g_pool_mu;
std::vector<Arena *> g_pool;  // idle arenas, any device

constexpr size_t kAlign = 256;
size_t aligned(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

struct Piece {
  const double *h;
  size_t n;  // doubles
};

struct Lease {
  Arena *a = nullptr;
  int next = 0;
  bool synced = true;  // no copy or kernel of this call can still be in flight

  int init() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(DFM_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    std::lock_guard<std::mutex> lock(g_pool_mu);
    for (size_t i = 0; i < g_pool.size(); i++)
      if (g_pool[i]->device == dev) {
        a = g_pool[i];
        g_pool.erase(g_pool.begin() + (long)i);
        return DFM_OK;
      }
    a = new Arena;
    a->device = dev;
    return DFM_OK;
  }
  // Slot `next` holding at least `bytes` (grown by 1/4 beyond the request, so slowly
  // growing calls reallocate seldom).
  int buf(size_t bytes, void **out) {
    *out = nullptr;
    if (bytes == 0) return DFM_OK;
    if (next >= Arena::kSlots) return set_error(DFM_ERR_ARG, "internal: arena slots exhausted");
    const int i = next++;
    if (a->cap[i] < bytes) {
      if (a->p[i]) {
        (void)hipFree(a->p[i]);
        a->p[i] = nullptr;
        a->cap[i] = 0;
      }
      const size_t want = bytes + bytes / 4;
      hipError_t e = hipMalloc(&a->p[i], want);
      if (e != hipSuccess) {
        a->p[i] = nullptr;
        return set_error(DFM_ERR_NOMEM, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
      }
      a->cap[i] = want;
    }
    *out = a->p[i];
    synced = false;
    return DFM_OK;
  }
  int staging(size_t bytes) {
    if (a->host_cap >= bytes) return DFM_OK;
    if (a->host) {
      // this call's input copy may still be reading the buffer about to be replaced
      (void)hipStreamSynchronize(hipStreamPerThread);
      (void)hipHostFree(a->host);
      a->host = a->host_dev = nullptr;
      a->host_cap = 0;
    }
    const size_t want = bytes + bytes / 4;
    hipError_t e = hipHostMalloc(&a->host, want, hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&a->host_dev, a->host, 0);
    if (e != hipSuccess) {
      if (a->host) (void)hipHostFree(a->host);
      a->host = a->host_dev = nullptr;
      return set_error(DFM_ERR_NOMEM, "hipHostMalloc(%zu): %s", want, hipGetErrorString(e));
    }
    a->host_cap = want;
    return DFM_OK;
  }
  // The inputs packed into one device buffer with one copy; d[i] receives piece i's device
  // address (nullptr for an empty piece).
  int upload(std::initializer_list<Piece> pieces, double **d) {
    size_t total = 0;
    for (const Piece &q : pieces) total += aligned(q.n * sizeof(double));
    void *dev = nullptr;
    int rc = buf(total, &dev);
    if (!rc) rc = staging(total);
    if (rc) return rc;
    size_t off = 0;
    int i = 0;
    for (const Piece &q : pieces) {
      d[i++] = q.n ? (double *)((char *)dev + off) : nullptr;
      if (q.n) std::memcpy((char *)a->host + off, q.h, q.n * sizeof(double));
      off += aligned(q.n * sizeof(double));
    }
    // the staging copies are copy kernels on the call's stream reading (or writing) the
    // mapped pinned buffer: small DMA copies from several threads' streams queue on the few
    // copy engines (8 threads: 84,000 calls/s against 28,000 with DMA copies)
    return total ? copy_d2d(hipStreamPerThread, dev, a->host_dev, total) : DFM_OK;
  }
  // Device buffers for the outputs, packed in one allocation (counts in doubles).
  int outputs(std::initializer_list<size_t> counts, double **d) {
    size_t total = 0;
    for (size_t n : counts) total += aligned(n * sizeof(double));
    void *dev = nullptr;
    int rc = buf(total, &dev);
    if (rc) return rc;
    size_t off = 0;
    int i = 0;
    for (size_t n : counts) {
      d[i++] = n ? (double *)((char *)dev + off) : nullptr;
      off += aligned(n * sizeof(double));
    }
    out_dev = dev;
    out_total = total;
    return DFM_OK;
  }
  // One copy of the packed outputs back, the call's single synchronisation, then the
  // outputs to the caller's arrays (a null host pointer skips that output).
  int download(std::initializer_list<double *> hosts, std::initializer_list<size_t> counts) {
    int rc = staging(out_total);
    if (rc) return rc;
    if (out_total) {
      rc = copy_d2d(hipStreamPerThread, a->host_dev, out_dev, out_total);
      if (rc) return rc;
    }
    hipError_t e = hipStreamSynchronize(hipStreamPerThread);
    if (e != hipSuccess) return set_error(DFM_ERR_HIP, "sync: %s", hipGetErrorString(e));
    synced = true;
    size_t off = 0;
    auto h = hosts.begin();
    for (size_t n : counts) {
      if (*h && n) std::memcpy(*h, (char *)a->host + off, n * sizeof(double));
      off += aligned(n * sizeof(double));
      ++h;
    }
    return DFM_OK;
  }
  ~Lease() {
    if (!a) return;
    // an early error return can leave this call's copies queued: let them finish before
    // another call reuses (or regrows) the buffers
    if (!synced) (void)hipStreamSynchronize(hipStreamPerThread);
    std::lock_guard<std::mutex> lock(g_pool_mu);
    g_pool.push_back(a);
  }

 private:
  void *out_dev = nullptr;
  size_t out_total = 0;
};

#define TRY(expr)               \
  do {                          \
    int rc_ = (expr);           \
    if (rc_ != DFM_OK) return rc_; \
  } while (0)

}  // namespace

extern "C" {

int dfm_gauss_conv(int D, int DV, double sigma, const double *x, const double *y,
                   const double *beta, double *gamma, int nx, int ny, int prec) {
  clear_error();
  TRY(check_common(D, DV, prec, sigma, nx, ny));
  if (nx == 0) return DFM_OK;
  if (!x || !gamma || (ny > 0 && (!y || !beta))) return set_error(DFM_ERR_ARG, "NULL argument");
  Lease L;
  TRY(L.init());
  double *in[3], *out[1];
  void *dw;
  TRY(L.upload({{x, (size_t)nx * D}, {y, (size_t)ny * D}, {beta, (size_t)ny * DV}}, in));
  TRY(L.outputs({(size_t)nx * DV}, out));
  size_t ws = dfm_workspace_size(DFM_OP_CONV, D, DV, nx, ny, prec);
  TRY(L.buf(ws, &dw));
  TRY(dfm_conv(hipStreamPerThread, D, DV, sigma, prec, in[0], nx, in[1], in[2], ny, out[0], dw,
               ws));
  return L.download({gamma}, {(size_t)nx * DV});
}

int dfm_gauss_grad1conv(int D, int DV, double sigma, const double *alpha, const double *x,
                        const double *y, const double *beta, double *gamma, int nx, int ny,
                        int prec) {
  clear_error();
  TRY(check_common(D, DV, prec, sigma, nx, ny));
  if (nx == 0) return DFM_OK;
  if (!x || !alpha || !gamma || (ny > 0 && (!y || !beta)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  Lease L;
  TRY(L.init());
  double *in[4], *out[1];
  void *dw;
  TRY(L.upload({{x, (size_t)nx * D}, {alpha, (size_t)nx * DV}, {y, (size_t)ny * D},
                {beta, (size_t)ny * DV}}, in));
  TRY(L.outputs({(size_t)nx * D}, out));
  size_t ws = dfm_workspace_size(DFM_OP_GRAD1, D, DV, nx, ny, prec);
  TRY(L.buf(ws, &dw));
  TRY(dfm_grad1(hipStreamPerThread, D, DV, sigma, prec, in[0], in[1], nx, in[2], in[3], ny,
                out[0], dw, ws));
  return L.download({gamma}, {(size_t)nx * D});
}

int dfm_gauss_gradconv(int D, int DV, double sigma, const double *x, const double *y,
                       const double *beta, unsigned dim, double *gamma, int nx, int ny,
                       int prec) {
  clear_error();
  TRY(check_common(D, DV, prec, sigma, nx, ny));
  if (dim >= (unsigned)D) return set_error(DFM_ERR_INDEX, "dimension index out of bounds");
  if (nx == 0) return DFM_OK;
  if (!x || !gamma || (ny > 0 && (!y || !beta))) return set_error(DFM_ERR_ARG, "NULL argument");
  Lease L;
  TRY(L.init());
  double *in[3], *out[1];
  void *dw;
  TRY(L.upload({{x, (size_t)nx * D}, {y, (size_t)ny * D}, {beta, (size_t)ny * DV}}, in));
  TRY(L.outputs({(size_t)nx * DV}, out));
  size_t ws = dfm_workspace_size(DFM_OP_GRADDIM, D, DV, nx, ny, prec);
  TRY(L.buf(ws, &dw));
  TRY(dfm_graddim(hipStreamPerThread, D, DV, sigma, prec, in[0], nx, in[1], in[2], ny, (int)dim,
                  out[0], dw, ws));
  return L.download({gamma}, {(size_t)nx * DV});
}

int dfm_gauss_gradconv_varlin(int D, int DV, double sigma, const double *x, const double *y,
                              const double *beta, double *gamma, int nx, int ny, int prec) {
  clear_error();
  TRY(check_common(D, DV, prec, sigma, nx, ny));
  if (nx == 0) return DFM_OK;
  if (!x || !gamma || (ny > 0 && (!y || !beta))) return set_error(DFM_ERR_ARG, "NULL argument");
  Lease L;
  TRY(L.init());
  double *in[3], *out[1];
  void *dw;
  TRY(L.upload({{x, (size_t)nx * D}, {y, (size_t)ny * D}, {beta, (size_t)ny * DV}}, in));
  TRY(L.outputs({(size_t)nx * DV * D}, out));
  size_t ws = dfm_workspace_size(DFM_OP_VARLIN, D, DV, nx, ny, prec);
  TRY(L.buf(ws, &dw));
  TRY(dfm_varlin(hipStreamPerThread, D, DV, sigma, prec, in[0], nx, in[1], in[2], ny, out[0], dw,
                 ws));
  return L.download({gamma}, {(size_t)nx * DV * D});
}

int dfm_gauss_hessianconv(int D, int DV, double sigma, const double *x, const double *y,
                          const double *beta, double *gamma, int nx, int ny, int prec) {
  clear_error();
  TRY(check_common(D, DV, prec, sigma, nx, ny));
  if (nx == 0) return DFM_OK;
  if (!x || !gamma || (ny > 0 && (!y || !beta))) return set_error(DFM_ERR_ARG, "NULL argument");
  Lease L;
  TRY(L.init());
  double *in[3], *out[1];
  void *dw;
  TRY(L.upload({{x, (size_t)nx * D}, {y, (size_t)ny * D}, {beta, (size_t)ny * DV}}, in));
  TRY(L.outputs({(size_t)nx * DV * D * D}, out));
  size_t ws = dfm_workspace_size(DFM_OP_HESSIAN, D, DV, nx, ny, prec);
  TRY(L.buf(ws, &dw));
  TRY(dfm_hessian(hipStreamPerThread, D, DV, sigma, prec, in[0], nx, in[1], in[2], ny, out[0], dw,
                  ws));
  return L.download({gamma}, {(size_t)nx * DV * D * D});
}

int dfm_gauss_graddiffconv(int D, int DV, double sigma, const double *x, const double *beta,
                           const double *eta, double *gamma, int nx, int prec) {
  clear_error();
  TRY(check_common(D, DV, prec, sigma, nx, nx));
  if (DV != D) return set_error(DFM_ERR_DIM, "Invalid number of columns of beta");
  if (nx == 0) return DFM_OK;
  if (!x || !beta || !eta || !gamma) return set_error(DFM_ERR_ARG, "NULL argument");
  Lease L;
  TRY(L.init());
  double *in[3], *out[1];
  void *dw;
  TRY(L.upload({{x, (size_t)nx * D}, {beta, (size_t)nx * DV}, {eta, (size_t)nx * D}}, in));
  TRY(L.outputs({(size_t)nx * D}, out));
  size_t ws = dfm_workspace_size(DFM_OP_GRADDIFF, D, DV, nx, nx, prec);
  TRY(L.buf(ws, &dw));
  TRY(dfm_graddiff(hipStreamPerThread, D, DV, sigma, prec, in[0], in[1], in[2], nx, out[0], dw,
                   ws));
  return L.download({gamma}, {(size_t)nx * D});
}

int dfm_gauss_conv_grad1(int D, int DV, double sigma, const double *x, const double *alpha,
                         const double *y, const double *beta, double *conv, double *grad, int nx,
                         int ny, int prec) {
  clear_error();
  TRY(check_common(D, DV, prec, sigma, nx, ny));
  if (nx == 0 || (!conv && !grad)) return DFM_OK;
  if (!x || (grad && !alpha) || (ny > 0 && (!y || !beta)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  Lease L;
  TRY(L.init());
  double *in[4], *out[2];
  void *dw;
  const size_t nc = conv ? (size_t)nx * DV : 0, ng = grad ? (size_t)nx * D : 0;
  TRY(L.upload({{x, (size_t)nx * D}, {alpha, grad ? (size_t)nx * DV : 0}, {y, (size_t)ny * D},
                {beta, (size_t)ny * DV}}, in));
  TRY(L.outputs({nc, ng}, out));
  const int op = conv && grad ? DFM_OP_CONV_GRAD1 : conv ? DFM_OP_CONV : DFM_OP_GRAD1;
  size_t ws = dfm_workspace_size(op, D, DV, nx, ny, prec);
  TRY(L.buf(ws, &dw));
  TRY(dfm_conv_grad1(hipStreamPerThread, D, DV, sigma, prec, in[0], in[1], nx, in[2], in[3], ny,
                     out[0], out[1], dw, ws));
  return L.download({conv, grad}, {nc, ng});
}


}  // extern "C"

// ---- ragged batches ---------------------------------------------------------------------

namespace {

struct BatchLayout {
  long *xoff, *yoff;
  void *pw;
  size_t pw_bytes, total;
  BatchLayout(char *base, int op, int prec, int D, int DV, const Seg &xs, const Seg &ys) {
    auto a256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
    size_t off = 0;
    xoff = base ? (long *)(base + off) : nullptr;
    off += a256(sizeof(long) * (xs.n + 1));
    yoff = base ? (long *)(base + off) : nullptr;
    off += a256(sizeof(long) * (ys.n + 1));
    pw_bytes = pairsum_batch_workspace(op, prec, D, DV, xs, ys);
    pw = base ? (void *)(base + off) : nullptr;
    off += a256(pw_bytes);
    total = off;
  }
};

int batch_op(int op_flags_conv, int op_flags_grad) {
  return op_flags_conv && op_flags_grad ? OP_CONVGRAD1 : op_flags_conv ? OP_CONV : OP_GRAD1;
}

}  // namespace

extern "C" {

size_t dfm_batch_workspace(int op, int dimpoint, int dimvect, int nseg, const long *x_offsets,
                           const long *y_offsets, int precision) {
  if (op != DFM_OP_CONV && op != DFM_OP_GRAD1 && op != DFM_OP_CONV_GRAD1) return 0;
  Seg xs, ys;
  if (seg_from_host(xs, nseg, x_offsets, nullptr) || seg_from_host(ys, nseg, y_offsets, nullptr))
    return 0;
  if (check_common(dimpoint, dimvect, precision, 1.0, xs.total, ys.total)) return 0;
  return BatchLayout(nullptr, op, precision, dimpoint, dimvect, xs, ys).total;
}

int dfm_batch_conv_grad1(void *stream, int D, int DV, double sigma, int prec, int nseg,
                         const long *x_offsets, const long *y_offsets, const double *x,
                         const double *alpha, const double *y, const double *beta, double *conv,
                         double *grad, void *work, size_t work_bytes) {
  clear_error();
  Seg xs, ys;
  int rc = seg_from_host(xs, nseg, x_offsets, nullptr);
  if (!rc) rc = seg_from_host(ys, nseg, y_offsets, nullptr);
  if (!rc) rc = check_common(D, DV, prec, sigma, xs.total, ys.total);
  if (!rc && xs.total > 0) rc = need(x, "x");
  if (!rc && xs.total > 0 && grad) rc = need(alpha, "alpha");
  if (!rc && xs.total > 0 && ys.total > 0) rc = need(y, "y") ? DFM_ERR_ARG : need(beta, "beta");
  if (rc) return rc;
  if (!conv && !grad) return DFM_OK;
  const int op = batch_op(conv != nullptr, grad != nullptr);
  BatchLayout need_(nullptr, op, prec, D, DV, xs, ys);
  if (!work || work_bytes < need_.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need_.total);
  BatchLayout L((char *)work, op, prec, D, DV, xs, ys);
  hipStream_t st = (hipStream_t)stream;
  if (nseg > 1) {
    if ((rc = put_longs(st, L.xoff, x_offsets, nseg + 1))) return rc;
    if ((rc = put_longs(st, L.yoff, y_offsets, nseg + 1))) return rc;
    xs.d = L.xoff;
    ys.d = L.yoff;
  }
  KParams p = blank();
  p.x = x; p.ta = alpha; p.y = y; p.w = beta;
  if (op == OP_GRAD1) {
    p.out0 = grad;
    return pairsum_batch(st, op, prec, D, DV, sigma, p, xs, ys, xs.total * D, 0, L.pw, L.pw_bytes);
  }
  p.out0 = conv;
  p.out1 = grad;
  return pairsum_batch(st, op, prec, D, DV, sigma, p, xs, ys, xs.total * DV,
                       grad ? xs.total * D : 0, L.pw, L.pw_bytes);
}

}  // extern "C"
