// integrators.hip — device-resident geodesic shooting, landmark flow, and the discrete
// adjoint (backward pass) of the shoot + flow scheme, recomputed from checkpoints; for one
// subject or for a ragged batch of subjects in the same launches.
//
// Forward (Diffeos.cxx:589-710), all state in HBM, every kernel sum one pair-sum launch:
//   shoot  z = (p, a):  F(z) = (K(p,p) a, -g(p, a)),  g_i = sum_j (a_i.a_j) grad_1 K(p_i, p_j)
//          Euler  z_{t+1} = z_t + dt F(z_t)                         (Diffeos::Shoot)
//          Heun   z* = z_t + dt F(z_t); z_{t+1} = z_t + dt/2 (F(z_t) + F(z*))
//          One F evaluation is one fused Convolve + ConvolveGradient(X, a) call with
//          targets = sources = p.
//   flow   v(x; z) = sum_j K(x, p_j) a_j;  Euler or Heun (improved Euler, the default)
//          (Diffeos::FlowLandmarkPointsTrajectory)
// Backward, form DFM_ADJOINT_DISCRETE: the exact transpose of those discrete steps:
//   J_F(z)^T (mu_p, mu_a) = (GradSym(mu_p) + SpecialHessian(mu_a), Convolve(mu_p) - Diff(mu_a))
//     one fused OP_SHOOTVJP pair sum
//   J_x v(x; z)^T th = Grad1(targets x, alpha th; sources p, weights a)
//   J_z v(x; z)^T th = (Grad1, Convolve)(targets p, alpha a; sources x, weights th)
//     one fused OP_CONVGRAD1 pair sum
// Backward, form DFM_ADJOINT_CONTINUOUS: the reference's continuous adjoint equations
// (AdjointEquationsIntegrator.cxx:59-136, 414-567), integrated backward along the forward
// states with the flow's Euler / improved-Euler choice:
//   d theta / ds = J_x v(x_s; z_s)^T theta                       (landmark cotangent)
//   d xi / ds    = J_F(z_s)^T xi + J_z v(x_s; z_s)^T theta_s      (ComputeUpdateAt)
// where ComputeUpdateAt's six terms are exactly J_F^T xi (dxi3 + dxi4, dxi5 + dxi6) and the
// landmark coupling (dxi1, dxi2), so the same two fused pair sums serve both forms.
// The forward stores (p, a, x) every k-th step and at the end; the backward recomputes the
// interior states of each segment of at most k steps from its checkpoint (the segment's
// last state is the next checkpoint) and walks it in reverse.  With k = 1 nothing is
// recomputed: the whole trajectory stays resident, (T + 1) x (2n + m) x d doubles.
//
// Batches (the per-subject loop of AbstractAtlas::ComputeDataTermGradient,
// AbstractAtlas.cxx:490-560): subject s owns control points [cp[s], cp[s + 1]) and
// landmarks [lm[s], lm[s + 1]) of concatenated arrays; trajectories are time-major,
// [T][all points][d].  Every pair sum of a step covers all subjects in one launch
// (grid.z = subject), every elementwise update is one launch over all subjects, and the
// zero-momenta exit is decided per subject.  Batched calls launch kernels only (no
// allocation, memset, host copy or synchronisation), so a caller may capture them in a
// HIP graph and replay it.
//
// Small single-subject shoots (at most kPersistMaxPoints control points, a stream that is
// not being captured) run as one cooperative launch for the whole trajectory
// (k_shoot_persistent): each workgroup owns a few control points, each of its lanes sums
// a strided slice of all the sources, a wavefront reduction finishes each point's field,
// the owning lane applies the step, and a grid barrier separates the steps.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "dfm_hip.h"
#include "dfm_internal.h"

namespace dfm {
namespace {


// out = ca a + cb b + cc c (b, c may be null) over segment blockIdx.y of off (in points of
// D coordinates; the flat range [0, n) without offsets).  With a gate (squared norm of the
// segment's initial momenta) below kZeroMomenta2, out = a, or 0 when gate_zero is set.
__global__ __launch_bounds__(256) void k_lin(double *out, const double *a, double ca,
                                             const double *b, double cb, const double *c,
                                             double cc, long n, const double *gate,
                                             int gate_zero, const long *off, int D) {
  const int seg = blockIdx.y;
  long lo = 0, hi = n;
  if (off) {
    lo = off[seg] * D;
    hi = off[seg + 1] * D;
  }
  const long i = lo + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hi) return;
  if (gate && gate[seg] < kZeroMomenta2) {
    out[i] = gate_zero ? 0.0 : a[i];
    return;
  }
  double v = ca * a[i];
  if (b) v += cb * b[i];
  if (c) v += cc * c[i];
  out[i] = v;
}

// out[s] = sum of a_i^2 over segment s (one workgroup per segment)
__global__ __launch_bounds__(256) void k_sumsq(const double *a, long n, const long *off, int D,
                                               double *out) {
  __shared__ double sh[256];
  const int seg = blockIdx.x;
  long lo = 0, hi = n;
  if (off) {
    lo = off[seg] * D;
    hi = off[seg + 1] * D;
  }
  double s = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += 256) s += a[i] * a[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[seg] = sh[0];
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Bump allocator over the caller's workspace; a dry run (base == nullptr) only sizes it.
struct Arena {
  char *base;
  size_t off = 0;
  explicit Arena(void *b) : base((char *)b) {}
  double *take(size_t count) {
    double *p = base ? (double *)(base + off) : nullptr;
    off += align256(count * sizeof(double));
    return p;
  }
  long *take_longs(size_t count) {
    long *p = base ? (long *)(base + off) : nullptr;
    off += align256(count * sizeof(long));
    return p;
  }
  void *take_bytes(size_t bytes) {
    void *p = base ? (void *)(base + off) : nullptr;
    off += align256(bytes);
    return p;
  }
};

struct Ctx {
  hipStream_t st;
  int D;
  double sigma;
  int prec;
  void *pw;      // pair-sum workspace
  size_t pw_bytes;
  bool batch;    // capture-safe batched pair sums (no Hilbert sort)
  int *done = nullptr;  // split-completion counters of the fused steps (KParams::done)
};

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// flat out = ca a + cb b + cc c over n doubles
int lin(const Ctx &cx, double *out, const double *a, double ca, const double *b, double cb,
        const double *c, double cc, long n) {
  if (n <= 0) return DFM_OK;
  hipLaunchKernelGGL(k_lin, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, cx.st, out, a, ca, b,
                     cb, c, cc, n, nullptr, 0, nullptr, cx.D);
  return check_launch("k_lin");
}

// the same, gated per segment of sg by the segment's initial momenta
int glin(const Ctx &cx, const Seg &sg, double *out, const double *a, double ca, const double *b,
         double cb, const double *c, double cc, const double *gate, int gate_zero) {
  if (sg.total <= 0) return DFM_OK;
  const bool segs = sg.n > 1;
  const long len = (segs ? sg.max : sg.total) * cx.D;
  if (len <= 0) return DFM_OK;
  hipLaunchKernelGGL(k_lin, dim3((unsigned)((len + 255) / 256), (unsigned)sg.n), dim3(256), 0,
                     cx.st, out, a, ca, b, cb, c, cc, sg.total * cx.D, gate, gate_zero,
                     segs ? sg.d : nullptr, cx.D);
  return check_launch("k_lin");
}

int copy(const Ctx &cx, double *out, const double *a, long n) {
  return lin(cx, out, a, 1.0, nullptr, 0.0, nullptr, 0.0, n);
}

int norm2(const Ctx &cx, const double *a, const Seg &sg, double *out) {
  hipLaunchKernelGGL(k_sumsq, dim3((unsigned)sg.n), dim3(256), 0, cx.st, a, sg.total * cx.D,
                     sg.n > 1 ? sg.d : nullptr, cx.D, out);
  return check_launch("k_sumsq");
}

KParams blank() {
  KParams p;
  std::memset(&p, 0, sizeof(p));
  return p;
}

int ps(const Ctx &cx, int op, KParams p, const Seg &xs, const Seg &ys, long o0, long o1) {
  if (cx.batch)
    return pairsum_batch(cx.st, op, cx.prec, cx.D, cx.D, cx.sigma, p, xs, ys, o0, o1, cx.pw,
                         cx.pw_bytes);
  p.nx = xs.total;
  p.ny = ys.total;
  return pairsum(cx.st, op, cx.prec, cx.D, cx.D, cx.sigma, p, o0, o1, cx.pw, cx.pw_bytes);
}

// Linear epilogue of one output of a pair sum (KParams o_*): the output takes
// cb base + cs S + ce extra instead of the sum S (a null base or extra is left out).
struct Epi {
  const double *base;
  double cb, cs;
  const double *extra;
  double ce;
};

void set_epi(KParams &p, int k, const Epi *e) {
  if (!e) return;
  if (!p.o_on) {  // the other output stays S
    p.o_on = 1;
    for (int j = 0; j < 2; j++) {
      p.o_base[j] = p.o_extra[j] = nullptr;
      p.o_cb[j] = p.o_ce[j] = 0.0;
      p.o_cs[j] = 1.0;
    }
  }
  p.o_base[k] = e->base;
  p.o_cb[k] = e->cb;
  p.o_cs[k] = e->cs;
  p.o_extra[k] = e->extra;
  p.o_ce[k] = e->ce;
}

// conv (nx x D) and/or grad (nx x D) of sum_j K(x_i, y_j) w_j with alpha (nx x D), each
// through its epilogue ec / eg when given
int conv_grad1(const Ctx &cx, const double *x, const double *alpha, const Seg &xs,
               const double *y, const double *w, const Seg &ys, double *conv, double *grad,
               const Epi *ec = nullptr, const Epi *eg = nullptr) {
  KParams p = blank();
  p.x = x; p.ta = alpha; p.y = y; p.w = w;
  const long n = xs.total * cx.D;
  if (conv && grad) {
    p.out0 = conv; p.out1 = grad;
    set_epi(p, 0, ec);
    set_epi(p, 1, eg);
    return ps(cx, OP_CONVGRAD1, p, xs, ys, n, n);
  }
  if (conv) {
    p.out0 = conv;
    set_epi(p, 0, ec);
    return ps(cx, OP_CONV, p, xs, ys, n, 0);
  }
  p.out0 = grad;
  set_epi(p, 0, eg);
  return ps(cx, OP_GRAD1, p, xs, ys, n, 0);
}

int shoot_vjp(const Ctx &cx, const double *p_, const double *a, const double *mp,
              const double *ma, const Seg &cs, double *jp, double *ja, const Epi *ep = nullptr,
              const Epi *ea = nullptr) {
  KParams p = blank();
  p.x = p_; p.ta = a; p.tb = mp; p.tc = ma;
  p.y = p_; p.w = a; p.se = mp; p.se2 = ma;
  p.out0 = jp; p.out1 = ja;
  set_epi(p, 0, ep);
  set_epi(p, 1, ea);
  return ps(cx, OP_SHOOTVJP, p, cs, cs, cs.total * cx.D, cs.total * cx.D);
}

size_t pair_ws(int prec, int D, const Seg &cs, const Seg &ls, bool batch) {
  size_t w = 0;
  if (batch) {
    w = std::max(w, pairsum_batch_workspace(OP_CONVGRAD1, prec, D, D, cs, cs));
    w = std::max(w, pairsum_batch_workspace(OP_SHOOTVJP, prec, D, D, cs, cs));
    if (ls.total > 0) {
      w = std::max(w, pairsum_batch_workspace(OP_CONV, prec, D, D, ls, cs));
      w = std::max(w, pairsum_batch_workspace(OP_GRAD1, prec, D, D, ls, cs));
      w = std::max(w, pairsum_batch_workspace(OP_CONVGRAD1, prec, D, D, cs, ls));
    }
    return w;
  }
  const long n = cs.total, m = ls.total;
  w = std::max(w, pairsum_workspace(OP_CONVGRAD1, prec, D, D, n, n));
  w = std::max(w, pairsum_workspace(OP_SHOOTVJP, prec, D, D, n, n));
  if (m > 0) {
    w = std::max(w, pairsum_workspace(OP_CONV, prec, D, D, m, n));
    w = std::max(w, pairsum_workspace(OP_GRAD1, prec, D, D, m, n));
    w = std::max(w, pairsum_workspace(OP_CONVGRAD1, prec, D, D, n, m));
  }
  return w;
}

// ---- forward steps ------------------------------------------------------------------

// Target tiles (workgroups along the targets, times segments) of a pair sum whose targets
// are segmented by s: at most this many.
long tile_bound(const Seg &s) { return s.n > 1 ? (s.max / 64 + 1) * s.n : s.total / 64 + 1; }

// The driver's counters: one split-completion counter per target tile of its pair sums
// (KParams::done; targets are the control points or the landmarks), which also hold the
// grid barrier of the one-launch shoot (kBarrierInts).
struct Counters {
  int *done;
  long n;
  void take(Arena &ar, const Seg &cs, const Seg &ls) {
    n = std::max<long>(std::max(tile_bound(cs), ls.total > 0 ? tile_bound(ls) : 0), 32 * 17);
    done = (int *)ar.take((size_t)(n + 1) / 2);
  }
};

// zero them (a fill kernel: capture-safe); each pair sum leaves them zero again
int clear_done(const Ctx &cx, const Counters &c) {
  fill(cx.st, (double *)c.done, (c.n + 1) / 2, 0.0);
  return check_launch("clear counters");
}

struct ShootTmp {
  double *dpos, *dmom, *ps, *as, *dpos2, *dmom2;
  void take(Arena &ar, long nd) {
    dpos = ar.take(nd); dmom = ar.take(nd); ps = ar.take(nd); as = ar.take(nd);
    dpos2 = ar.take(nd); dmom2 = ar.take(nd);
  }
};

// One pair sum over the control points (p, a) with the step applied in the same launch
// (KParams e_*): p1 = pb [+ cx px] + c conv, a1 = ab [- cx ax] - c grad, gated by `gate`;
// conv and grad themselves go to dpos and dmom.
int fused_step_sum(const Ctx &cx, const double *p, const double *a, const Seg &cs, double *dpos,
                   double *dmom, double *p1, double *a1, const double *pb, const double *ab,
                   double c, const double *px, const double *ax, double cxx, const double *gate) {
  KParams q = blank();
  q.x = p; q.ta = a; q.y = p; q.w = a;
  q.out0 = dpos; q.out1 = dmom;
  q.e_on = 1;
  q.e_p1 = p1; q.e_a1 = a1; q.e_pb = pb; q.e_ab = ab;
  q.e_px = px; q.e_ax = ax; q.e_c = c; q.e_cx = cxx;
  q.e_gate = gate;
  // The step's own sums may finish their splits in the kernel (on grids of few target
  // tiles, where it saves launches; launch_timed decides).  Other sums keep the separate
  // reduction: the device-scope fence before each wavefront's arrival writes back the L2,
  // which on the large grids of the flow and the adjoint cost more than the launch it saves
  // (registration gradient 25.3 against 12.7 ms, same box).
  q.done = cx.done;
  const long nd = cs.total * cx.D;
  return ps(cx, OP_CONVGRAD1, q, cs, cs, nd, nd);
}

// (p, a) -> (p1, a1) by one Euler or Heun step: one pair sum per field evaluation, the
// update applied in its kernel on small grids, else after its split reduction.
int shoot_step(const Ctx &cx, int scheme, const double *p, const double *a, const Seg &cs,
               double dt, double *p1, double *a1, const ShootTmp &t, const double *gate) {
  if (scheme == DFM_EULER)
    return fused_step_sum(cx, p, a, cs, t.dpos, t.dmom, p1, a1, p, a, dt, nullptr, nullptr, 0.0,
                          gate);
  // Heun: predictor (ungated), then the corrector from the original state
  int rc = fused_step_sum(cx, p, a, cs, t.dpos, t.dmom, t.ps, t.as, p, a, dt, nullptr, nullptr,
                          0.0, nullptr);
  if (rc) return rc;
  return fused_step_sum(cx, t.ps, t.as, cs, t.dpos2, t.dmom2, p1, a1, p, a, 0.5 * dt, t.dpos,
                        t.dmom, 0.5 * dt, gate);
}

struct FlowTmp {
  double *v1, *v2, *xs;
  void take(Arena &ar, long md) { v1 = ar.take(md); v2 = ar.take(md); xs = ar.take(md); }
};

// x -> x1 in the flow from (p, a) to (p1, a1).
int flow_step(const Ctx &cx, int improved, const double *p, const double *a, const double *p1,
              const double *a1, const Seg &cs, const double *x, const Seg &ls, double dt,
              double *x1, const FlowTmp &t, const double *gate) {
  int rc;
  if (!improved) {
    if ((rc = conv_grad1(cx, x, nullptr, ls, p, a, cs, t.v1, nullptr))) return rc;
    return glin(cx, ls, x1, x, 1.0, t.v1, dt, nullptr, 0.0, gate, 0);
  }
  // xs = x + dt v(x; z) from the first sum's epilogue; x1 = x + dt/2 (v1 + v2) as
  // (x + xs)/2 + dt/2 v2, gated
  const Epi e{x, 1.0, dt, nullptr, 0.0};
  if ((rc = conv_grad1(cx, x, nullptr, ls, p, a, cs, t.xs, nullptr, &e))) return rc;
  if ((rc = conv_grad1(cx, t.xs, nullptr, ls, p1, a1, cs, t.v2, nullptr))) return rc;
  return glin(cx, ls, x1, x, 0.5, t.xs, 0.5, t.v2, 0.5 * dt, gate, 0);
}

int check_common(int D, int prec, double sigma, long n, long m, int T) {
  if (D != 2 && D != 3) return set_error(DFM_ERR_DIM, "unsupported dimpoint %d (2 or 3)", D);
  if (!valid_precision(prec))
    return set_error(DFM_ERR_ARG, "unknown precision %d", prec);
  if (!(sigma == sigma)) return set_error(DFM_ERR_ARG, "kernel width is NaN");
  if (n < 0 || m < 0 || n > 0x7fffffffL || m > 0x7fffffffL)
    return set_error(DFM_ERR_ARG, "point counts must be in [0, 2^31)");
  if (T < 1) return set_error(DFM_ERR_ARG, "number of time points must be >= 1");
  return DFM_OK;
}

// ---- layouts (shared by the workspace queries and the drivers) ----------------------

// device copies of the segment offsets and the per-segment zero-momenta gates
struct SegArrays {
  long *cp, *lm;
  double *gate;
  void take(Arena &ar, int nseg) {
    cp = ar.take_longs(nseg + 1);
    lm = ar.take_longs(nseg + 1);
    gate = ar.take(nseg);
  }
};

struct ShootLayout {
  SegArrays sa;
  ShootTmp t;
  Counters c;
  void *pw;
  size_t pw_bytes;
  size_t total;
  ShootLayout(void *base, int prec, int D, const Seg &cs, bool batch) {
    Arena ar(base);
    sa.take(ar, cs.n);
    t.take(ar, cs.total * D);
    c.take(ar, cs, Seg::single(0));
    pw_bytes = pair_ws(prec, D, cs, Seg::single(0), batch);
    pw = ar.take_bytes(pw_bytes);
    total = ar.off;
  }
};

struct FlowLayout {
  SegArrays sa;
  FlowTmp t;
  Counters c;
  void *pw;
  size_t pw_bytes;
  size_t total;
  FlowLayout(void *base, int prec, int D, const Seg &cs, const Seg &ls, bool batch) {
    Arena ar(base);
    sa.take(ar, cs.n);
    t.take(ar, ls.total * D);
    c.take(ar, cs, ls);
    pw_bytes = pair_ws(prec, D, cs, ls, batch);
    pw = ar.take_bytes(pw_bytes);
    total = ar.off;
  }
};

struct AdjointLayout {
  int segs, k;
  SegArrays sa;
  double *ckpt;   // (segs + 1) x (2n + m) x D: states 0, k, 2k, ... and the final one
  double *seg;    // (k + 1) x (2n + m) x D
  ShootTmp st;
  FlowTmp ft;
  Counters c;
  double *lp, *la, *lpn, *lan, *lpf, *laf, *jp, *ja, *jsp, *jsa, *nup, *nua;  // n x D
  double *th, *thn, *ths, *mu;                                               // m x D
  void *pw;
  size_t pw_bytes;
  size_t total;
  AdjointLayout(void *base, int prec, int D, const Seg &cs, const Seg &ls, int T, int kk,
                bool batch) {
    Arena ar(base);
    const long n = cs.total, m = ls.total;
    k = std::max(1, kk);
    segs = T > 1 ? (T - 1 + k - 1) / k : 0;
    const long rec = (2 * n + m) * D;
    sa.take(ar, cs.n);
    ckpt = ar.take((size_t)(segs + 1) * rec);
    seg = ar.take((size_t)(std::min(k, std::max(T - 1, 0)) + 1) * rec);
    st.take(ar, n * D);
    ft.take(ar, m * D);
    c.take(ar, cs, ls);
    double **vn[] = {&lp, &la, &lpn, &lan, &lpf, &laf, &jp, &ja, &jsp, &jsa, &nup, &nua};
    for (double **v : vn) *v = ar.take(n * D);
    double **vm[] = {&th, &thn, &ths, &mu};
    for (double **v : vm) *v = ar.take(m * D);
    pw_bytes = pair_ws(prec, D, cs, ls, batch);
    pw = ar.take_bytes(pw_bytes);
    total = ar.off;
  }
};

// The checkpoint interval of a call: k >= 1 as given; k = 0 keeps the whole trajectory,
// (T + 1) x (2n + m) x D doubles, when it fits kResidentTrajectoryBytes (HBM is plentiful:
// no backward recompute), else every kAutoCheckpoint-th state.  Results do not depend on it.
int resolve_checkpoint(int k, int T, long n, long m, int D) {
  if (k != 0) return k;
  const double bytes = (double)(T + 1) * (double)(2 * n + m) * D * sizeof(double);
  return bytes <= kResidentTrajectoryBytes ? 1 : kAutoCheckpoint;
}

// segment descriptors of a call (nseg == 0: one subject with plain counts n, m)
int segs_of(int nseg, const long *cp_h, const long *lm_h, long n, long m, Seg &cs, Seg &ls) {
  if (nseg == 0) {
    cs = Seg::single(n);
    ls = Seg::single(m);
    return DFM_OK;
  }
  int rc = seg_from_host(cs, nseg, cp_h, nullptr);
  if (rc) return rc;
  if (lm_h) return seg_from_host(ls, nseg, lm_h, nullptr);
  ls = Seg::single(0);
  ls.n = nseg;
  return DFM_OK;
}

// device copies of the offsets, written into the workspace by kernels
int upload_segs(const Ctx &cx, Seg &cs, Seg &ls, const SegArrays &sa) {
  if (cs.n <= 1) return DFM_OK;
  int rc = put_longs(cx.st, sa.cp, cs.h, cs.n + 1);
  if (rc) return rc;
  cs.d = sa.cp;
  if (ls.h) {
    if ((rc = put_longs(cx.st, sa.lm, ls.h, ls.n + 1))) return rc;
    ls.d = sa.lm;
  }
  return DFM_OK;
}

// ---- one launch for a whole small shoot ------------------------------------------------

// A 1,000-point shoot (10 Euler steps, FP32C): 0.150 ms in one launch against 0.214 ms in
// ten fused-step launches; 2,048 points: 0.237 against 0.467 ms; 4,096 points: 0.421
// against 0.579 ms (kernel times).  At 1,000 points the launch with the field sums skipped
// takes 0.094 ms (launch, the barriers and the state writes) and the launch with the grid
// barrier skipped 0.102 ms: about 8 us per step of barrier and 8 of staging and sums.  A flat
// counter (one address, no tree) measured 0.172 ms, 64 lanes per point 0.193, 8 lanes 0.179.
constexpr long kPersistMaxPoints = 4096;
// lanes per point up to 1,024 points, then 16 (about 256 workgroups either way)
constexpr int kPersistSpt = 32;
// 0: cooperative-groups grid sync; 1: the counter barrier below
constexpr int kPersistBarrier = 1;

struct PersistShoot {
  unsigned *count;              // barrier counter, zero at launch
  const double *p0, *a0;
  double *P, *A;                // trajectories, T x n x D
  double *ps, *as, *dp, *da;    // Heun: predicted state and the first field, n x D
  long n;
  int T, scheme;
  int spt;                      // lanes (source slices) per control point: 16, 32 or 64
  double dt;
  double inv_s2;                // 1 / sigma^2
  double cg;                    // -2 / sigma^2
};

// Sources staged through LDS per chunk of this many points (positions and momenta, FP64)
constexpr int kPersistChunk = 1024;

// Field F = (sum_j K_ij a_j, -2/sigma^2 sum_j K_ij (a_i.a_j)(x_i - x_j)) of point i over the
// sources j = slice, slice + spt, ... of state (p, m), staged through LDS (`sp`, `sm`) one
// chunk at a time by the whole workgroup; then the sum over the spt lanes that share point i
// (lane bits below spt, xor butterfly: the same order in every lane).
// F32: the kernel value exp2 in FP32 of the FP64 squared distance (the DFM_FP32C kernel
// form; FP64 differences and accumulation cost nothing at this size).
template <bool F32, int D>
__device__ __forceinline__ void persist_field(const PersistShoot &a, const double *p,
                                              const double *m, long i, int slice,
                                              double (&conv)[D], double (&grad)[D],
                                              double *sp, double *sm) {
  double xi[D], ai[D];
#pragma unroll
  for (int k = 0; k < D; k++) {
    conv[k] = 0.0;
    grad[k] = 0.0;
    xi[k] = i < a.n ? p[i * D + k] : 0.0;
    ai[k] = i < a.n ? m[i * D + k] : 0.0;
  }
  for (long c0 = 0; c0 < a.n; c0 += kPersistChunk) {
    const int cn = (int)std::min<long>(kPersistChunk, a.n - c0);
    __syncthreads();  // the previous chunk is consumed
    const double *pc = p + c0 * D, *mc = m + c0 * D;
    if ((((uintptr_t)pc | (uintptr_t)mc) & 15) == 0 && (cn * D) % 2 == 0) {
      // 16-byte loads, eight in flight per lane: 1,000 points stage in 6 rounds instead of
      // 12 (1,000-point shoot 0.150 against 0.166 ms, 4,096 points 0.421 against 0.480 ms)
      const double2 *p2 = reinterpret_cast<const double2 *>(pc);
      const double2 *m2 = reinterpret_cast<const double2 *>(mc);
#pragma unroll 8
      for (int q = threadIdx.x; q < cn * D / 2; q += 256) {
        const double2 u = p2[q], v = m2[q];
        sp[2 * q] = u.x;
        sp[2 * q + 1] = u.y;
        sm[2 * q] = v.x;
        sm[2 * q + 1] = v.y;
      }
    } else {
      for (int q = threadIdx.x; q < cn * D; q += 256) {
        sp[q] = pc[q];
        sm[q] = mc[q];
      }
    }
    __syncthreads();
    if (i >= a.n) continue;
#pragma unroll 4
    for (int j = slice; j < cn; j += a.spt) {
      double d[D], aj[D], r2 = 0.0, aa = 0.0;
#pragma unroll
      for (int k = 0; k < D; k++) {
        d[k] = xi[k] - sp[j * D + k];
        aj[k] = sm[j * D + k];
        r2 += d[k] * d[k];
        aa += ai[k] * aj[k];
      }
      double K;
      if constexpr (F32)
        K = (double)__builtin_amdgcn_exp2f((float)(-r2 * a.inv_s2 * 1.4426950408889634074));
      else
        K = exp(-r2 * a.inv_s2);
      const double t = K * aa;
#pragma unroll
      for (int k = 0; k < D; k++) {
        conv[k] += K * aj[k];
        grad[k] += t * d[k];
      }
    }
  }
  for (int off = a.spt / 2; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < D; k++) {
      conv[k] += __shfl_xor(conv[k], off);
      grad[k] += __shfl_xor(grad[k], off);
    }
#pragma unroll
  for (int k = 0; k < D; k++) grad[k] *= a.cg;
}

// Grid barrier number `k` (1, 2, ...) of a cooperative launch (every workgroup resident),
// as a two-level tree of counters, each on its own 128-byte line (count[32 (1 + l)] for
// leaf l, count[0] the root): a workgroup arrives at leaf blockIdx.x % kBarrierLeaves, the
// last arrival of a leaf arrives at the root, and everyone waits for the root.  Arrivals on
// one address are serialised by the memory, so the tree splits them over the leaves.  The
// counters only grow (zero at launch); the wait is bounded (about a second), so the launch
// always ends.
constexpr int kBarrierLeaves = 16;
constexpr int kBarrierInts = 32 * (1 + kBarrierLeaves);
__device__ __forceinline__ void counter_barrier(unsigned *count, unsigned k) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned g = gridDim.x;
    const unsigned leaf = blockIdx.x % kBarrierLeaves;
    const unsigned leaves = g < (unsigned)kBarrierLeaves ? g : (unsigned)kBarrierLeaves;
    const unsigned leaf_size = (g - leaf + kBarrierLeaves - 1) / kBarrierLeaves;
    __threadfence();  // this workgroup's writes before its arrival
    if (atomicAdd(&count[32 * (1 + leaf)], 1u) + 1 == k * leaf_size) atomicAdd(&count[0], 1u);
    const unsigned target = k * leaves;
    for (unsigned spins = 0; spins < (1u << 24); spins++) {
      if (__hip_atomic_load(&count[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target)
        break;
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();  // the other workgroups' writes after everyone's arrival
  }
  __syncthreads();
}

template <bool F32, int D>
__global__ __launch_bounds__(256) void k_shoot_persistent(PersistShoot a) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  unsigned barriers = 0;
  auto sync = [&]() {
    if constexpr (kPersistBarrier == 1)
      counter_barrier(a.count, ++barriers);
    else
      grid.sync();
  };
  __shared__ double sred[256];
  __shared__ double sp[kPersistChunk * D], sm[kPersistChunk * D];
  const int tid = threadIdx.x;
  const long i = (long)blockIdx.x * (256 / a.spt) + tid / a.spt;  // this lane's point
  const int slice = tid % a.spt;
  const bool owner = slice == 0 && i < a.n;
  const long nd = a.n * D;

  // |a0|^2 (every workgroup sums it: no barrier needed before the first step)
  double s = 0.0;
  for (long k = tid; k < nd; k += 256) s += a.a0[k] * a.a0[k];
  sred[tid] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) sred[tid] += sred[tid + st];
    __syncthreads();
  }
  const bool frozen = sred[0] < kZeroMomenta2;  // Diffeos.cxx:609-611
  if (owner)
#pragma unroll
    for (int k = 0; k < D; k++) {
      a.P[i * D + k] = a.p0[i * D + k];
      a.A[i * D + k] = a.a0[i * D + k];
    }

  double conv[D], grad[D];
  for (int t = 0; t + 1 < a.T; t++) {
    const double *p = t == 0 ? a.p0 : a.P + t * nd;
    const double *m = t == 0 ? a.a0 : a.A + t * nd;
    double *p1 = a.P + (t + 1) * nd;
    double *m1 = a.A + (t + 1) * nd;
    persist_field<F32, D>(a, p, m, i, slice, conv, grad, sp, sm);
    if (a.scheme == DFM_EULER) {
      if (owner)
#pragma unroll
        for (int k = 0; k < D; k++) {
          const long q = i * D + k;
          double vp = 1.0 * p[q], va = 1.0 * m[q];
          vp += a.dt * conv[k];
          va += -a.dt * grad[k];
          p1[q] = frozen ? p[q] : vp;
          m1[q] = frozen ? 0.0 : va;
        }
    } else {  // Heun: predictor, barrier, corrector from the original state
      if (owner)
#pragma unroll
        for (int k = 0; k < D; k++) {
          const long q = i * D + k;
          double vp = 1.0 * p[q], va = 1.0 * m[q];
          vp += a.dt * conv[k];
          va += -a.dt * grad[k];
          a.ps[q] = vp;
          a.as[q] = va;
          a.dp[q] = conv[k];
          a.da[q] = grad[k];
        }
      sync();
      persist_field<F32, D>(a, a.ps, a.as, i, slice, conv, grad, sp, sm);
      if (owner)
#pragma unroll
        for (int k = 0; k < D; k++) {
          const long q = i * D + k;
          double vp = 1.0 * p[q], va = 1.0 * m[q];
          vp += 0.5 * a.dt * a.dp[q];
          va += -0.5 * a.dt * a.da[q];
          vp += 0.5 * a.dt * conv[k];
          va += -0.5 * a.dt * grad[k];
          p1[q] = frozen ? p[q] : vp;
          m1[q] = frozen ? 0.0 : va;
        }
    }
    if (t + 2 < a.T) sync();  // the next step reads every point of this one
  }
}

// The whole shoot in one cooperative launch when it fits (returns 1 if it ran, 0 to fall
// back to the per-step launches, or an error).
int shoot_persistent(hipStream_t st, int D, double sigma, int prec, int scheme, const double *p0,
                     const double *a0, long n, int T, double dt, double *P, double *A,
                     const ShootTmp &tmp, int *counters) {
  if (n > kPersistMaxPoints || T < 2) return 0;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return 0;
  }
  int dev = 0, coop = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      !coop) {
    (void)hipGetLastError();
    return 0;
  }
  PersistShoot a;
  a.count = (unsigned *)counters;
  a.p0 = p0; a.a0 = a0; a.P = P; a.A = A;
  a.ps = tmp.ps; a.as = tmp.as; a.dp = tmp.dpos; a.da = tmp.dmom;
  a.n = n; a.T = T; a.scheme = scheme; a.dt = dt;
  a.inv_s2 = 1.0 / (sigma * sigma);
  a.cg = -2.0 / (sigma * sigma);
  a.spt = n <= 1024 ? kPersistSpt : 16;
  const long grid = (n + 256 / a.spt - 1) / (256 / a.spt);
  const bool f32 = prec != DFM_FP64;
  const void *fn = D == 2 ? (f32 ? (const void *)k_shoot_persistent<true, 2>
                                 : (const void *)k_shoot_persistent<false, 2>)
                          : (f32 ? (const void *)k_shoot_persistent<true, 3>
                                 : (const void *)k_shoot_persistent<false, 3>);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess ||
      (long)per_cu * cus < grid) {
    (void)hipGetLastError();
    return 0;
  }
  static_assert(kBarrierInts <= 32 * 17, "barrier counters fit the counter array");
  fill(st, (double *)counters, kBarrierInts / 2, 0.0);  // the barrier counters
  void *args[] = {&a};
  hipEvent_t ev = profile_begin(st);
  hipError_t e = hipLaunchCooperativeKernel(fn, dim3((unsigned)grid), dim3(256), args, 0, st);
  // one fused Convolve + Grad1 over n x n pairs per field evaluation
  const double pairs = (double)n * n * (T - 1) * (scheme == DFM_HEUN ? 2 : 1);
  profile_end(st, ev, pairs, pairs * flops_per_pair(OP_CONVGRAD1, D, D));
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "cooperative shoot: %s", hipGetErrorString(e));
  return 1;
}

// ---- drivers (nseg == 0: one subject through the sorted single-call pair sums) ---------

int shoot_impl(void *stream, int D, double sigma, int prec, int scheme, int nseg,
               const long *cp_off, const double *p0, const double *a0, long n, int T,
               double t0, double tN, double *P, double *A, void *work, size_t work_bytes) {
  clear_error();
  Seg cs, ls;
  int rc = segs_of(nseg, cp_off, nullptr, n, 0, cs, ls);
  if (rc) return rc;
  if ((rc = check_common(D, prec, sigma, cs.total, 0, T))) return rc;
  if (scheme != DFM_EULER && scheme != DFM_HEUN) return set_error(DFM_ERR_ARG, "unknown scheme %d", scheme);
  if (cs.total == 0) return DFM_OK;
  if (!p0 || !a0 || !P || !A) return set_error(DFM_ERR_ARG, "NULL argument");
  const bool batch = nseg > 0;
  ShootLayout need(nullptr, prec, D, cs, batch);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  ShootLayout L(work, prec, D, cs, batch);
  Ctx cx{(hipStream_t)stream, D, sigma, prec, L.pw, L.pw_bytes, batch};
  const double dt = T > 1 ? (tN - t0) / (T - 1) : 0.0;
  if (!batch) {
    rc = shoot_persistent(cx.st, D, sigma, prec, scheme, p0, a0, cs.total, T, dt, P, A, L.t,
                          L.c.done);
    if (rc == 1) return DFM_OK;
    if (rc) return rc;
  }
  if ((rc = upload_segs(cx, cs, ls, L.sa)) || (rc = clear_done(cx, L.c))) return rc;
  cx.done = L.c.done;
  const long nd = cs.total * D;
  if ((rc = copy(cx, P, p0, nd)) || (rc = copy(cx, A, a0, nd)) || (rc = norm2(cx, a0, cs, L.sa.gate)))
    return rc;
  for (int t = 0; t + 1 < T; t++)
    if ((rc = shoot_step(cx, scheme, P + t * nd, A + t * nd, cs, dt, P + (t + 1) * nd,
                         A + (t + 1) * nd, L.t, L.sa.gate)))
      return rc;
  return DFM_OK;
}

int flow_impl(void *stream, int D, double sigma, int prec, int improved_euler, int nseg,
              const long *cp_off, const long *lm_off, const double *P, const double *A, long n,
              const double *x0, long m, int T, double t0, double tN, double *X, void *work,
              size_t work_bytes) {
  clear_error();
  if (nseg > 0 && !lm_off) return set_error(DFM_ERR_ARG, "landmark offsets are NULL");
  Seg cs, ls;
  int rc = segs_of(nseg, cp_off, lm_off, n, m, cs, ls);
  if (rc) return rc;
  if ((rc = check_common(D, prec, sigma, cs.total, ls.total, T))) return rc;
  if (ls.total == 0) return DFM_OK;
  if (!x0 || !X || (cs.total > 0 && (!P || !A))) return set_error(DFM_ERR_ARG, "NULL argument");
  const bool batch = nseg > 0;
  FlowLayout need(nullptr, prec, D, cs, ls, batch);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  FlowLayout L(work, prec, D, cs, ls, batch);
  Ctx cx{(hipStream_t)stream, D, sigma, prec, L.pw, L.pw_bytes, batch};
  if ((rc = upload_segs(cx, cs, ls, L.sa)) || (rc = clear_done(cx, L.c))) return rc;
  cx.done = L.c.done;
  const long nd = cs.total * D, md = ls.total * D;
  if ((rc = copy(cx, X, x0, md))) return rc;
  if (cs.total == 0) {
    for (int t = 1; t < T; t++)
      if ((rc = copy(cx, X + t * md, x0, md))) return rc;
    return DFM_OK;
  }
  if ((rc = norm2(cx, A, cs, L.sa.gate))) return rc;
  const double dt = T > 1 ? (tN - t0) / (T - 1) : 0.0;
  for (int t = 0; t + 1 < T; t++)
    if ((rc = flow_step(cx, improved_euler, P + t * nd, A + t * nd, P + (t + 1) * nd,
                        A + (t + 1) * nd, cs, X + t * md, ls, dt, X + (t + 1) * md, L.t,
                        L.sa.gate)))
      return rc;
  return DFM_OK;
}

// d xi at state z = (p, a, x) with landmark cotangent th: J_F(z)^T (xp, xm) (one SHOOTVJP
// pair sum) plus J_z v(x; z)^T th (one CONVGRAD1 pair sum: Grad1 onto p, Convolve onto a),
// written to (op, oa); lpf / laf are scratch.
int continuous_rate(const Ctx &cx, AdjointLayout &L, const Seg &cs, const Seg &ls,
                    const double *p, const double *a, const double *x, const double *th,
                    const double *xp, const double *xm, double *op, double *oa) {
  const long nd = cs.total * cx.D;
  int rc = shoot_vjp(cx, p, a, xp, xm, cs, op, oa);
  if (rc || ls.total == 0) return rc;
  if ((rc = conv_grad1(cx, p, a, cs, x, th, ls, L.laf, L.lpf))) return rc;
  if ((rc = lin(cx, op, op, 1.0, L.lpf, 1.0, nullptr, 0.0, nd))) return rc;
  return lin(cx, oa, oa, 1.0, L.laf, 1.0, nullptr, 0.0, nd);
}

// One backward step of the continuous adjoint equations from state t+1 (z1) to state t
// (zt): theta in L.th and xi in (L.lp, L.la) move from t+1 to t (AEI.cxx:96-133, 438-462).
// A jump condition src (m x d, or null) is added to both rates of theta, as the reference
// adds a regression observation's condition (AEI.cxx:104-110, 123-129).
int continuous_step(const Ctx &cx, AdjointLayout &L, const Seg &cs, const Seg &ls, int improved,
                    double dt, const double *pt, const double *at, const double *xt,
                    const double *p1, const double *a1, const double *x1,
                    const double *src = nullptr) {
  const long nd = cs.total * cx.D, md = ls.total * cx.D;
  int rc;
  if (md > 0) {
    // theta_t = theta_{t+1} + dt Grad1(x_{t+1}, theta_{t+1}; z_{t+1})  [+ Heun corrector]
    if ((rc = conv_grad1(cx, x1, L.th, ls, p1, a1, cs, nullptr, L.ths))) return rc;
    if (src && (rc = lin(cx, L.ths, L.ths, 1.0, src, 1.0, nullptr, 0.0, md))) return rc;
    if ((rc = lin(cx, L.thn, L.th, 1.0, L.ths, dt, nullptr, 0.0, md))) return rc;
    if (improved) {
      if ((rc = conv_grad1(cx, xt, L.thn, ls, pt, at, cs, nullptr, L.mu))) return rc;
      if (src && (rc = lin(cx, L.mu, L.mu, 1.0, src, 1.0, nullptr, 0.0, md))) return rc;
      if ((rc = lin(cx, L.thn, L.th, 1.0, L.ths, 0.5 * dt, L.mu, 0.5 * dt, md))) return rc;
    }
  }
  // xi_t = xi_{t+1} + dt rate(z_{t+1}, theta_{t+1}, xi_{t+1})  [+ Heun corrector at z_t]
  if ((rc = continuous_rate(cx, L, cs, ls, p1, a1, x1, L.th, L.lp, L.la, L.jp, L.ja))) return rc;
  if ((rc = lin(cx, L.lpn, L.lp, 1.0, L.jp, dt, nullptr, 0.0, nd))) return rc;
  if ((rc = lin(cx, L.lan, L.la, 1.0, L.ja, dt, nullptr, 0.0, nd))) return rc;
  if (improved) {
    if ((rc = continuous_rate(cx, L, cs, ls, pt, at, xt, L.thn, L.lpn, L.lan, L.jsp, L.jsa)))
      return rc;
    if ((rc = lin(cx, L.lpn, L.lp, 1.0, L.jp, 0.5 * dt, L.jsp, 0.5 * dt, nd))) return rc;
    if ((rc = lin(cx, L.lan, L.la, 1.0, L.ja, 0.5 * dt, L.jsa, 0.5 * dt, nd))) return rc;
  }
  std::swap(L.lp, L.lpn);
  std::swap(L.la, L.lan);
  if (md > 0) std::swap(L.th, L.thn);
  return DFM_OK;
}

// Jump conditions of the continuous adjoint (AEI.cxx:59-136): theta starts from the
// condition of the last jump if that jump is at T-1 (else from zero), and the condition at
// jump time t is added to theta's rate over the step t -> t-1.  The reference walks its
// list index down only in the improved-Euler branch, so under standard Euler at most one
// jump after the terminal one is applied; one condition alone is no jump (it starts theta
// whatever its time), and a jump at time 0 is never reached.  terminal = -1 for zero;
// step_src[t] = condition index of step t -> t-1, or -1.
struct JumpPlan {
  int terminal = 0;
  std::vector<int> step_src;
};

JumpPlan plan_jumps(int njumps, const int *times, int T, int improved_euler) {
  JumpPlan jp;
  jp.step_src.assign((size_t)std::max(T, 1), -1);
  if (njumps <= 1) return jp;
  int subj = njumps - 1;
  jp.terminal = -1;
  if (times[subj] == T - 1) {
    jp.terminal = subj;
    subj = std::max(subj - 1, 0);
  }
  for (int t = T - 1; t > 0; t--)
    if (t == times[subj]) {
      jp.step_src[t] = subj;
      if (improved_euler) subj = std::max(subj - 1, 0);
    }
  return jp;
}

// phase 1: forward pass only (checkpoints and the final landmarks, *x_final, stay in the
// workspace); phase 2: backward pass from the checkpoints of an earlier phase 1 with the
// same arguments; phase 3: both.  jumps: the continuous form's jump plan over the
// conditions in theta_T (njumps x m x d), or null for one condition at T-1.
int adjoint_impl(void *stream, int D, double sigma, int prec, int scheme, int improved_euler,
                 int nseg, const long *cp_off, const long *lm_off, const double *p0,
                 const double *a0, long n_, const double *x0, long m_, int T, double t0,
                 double tN, int checkpoint_every, const double *theta_T, double *grad_p0,
                 double *grad_a0, double *grad_x0, void *work, size_t work_bytes, int phase = 3,
                 const double **x_final = nullptr, int form = DFM_ADJOINT_DISCRETE,
                 const JumpPlan *jumps = nullptr) {
  clear_error();
  if (form != DFM_ADJOINT_DISCRETE && form != DFM_ADJOINT_CONTINUOUS)
    return set_error(DFM_ERR_ARG, "unknown adjoint form %d", form);
  if (jumps && form != DFM_ADJOINT_CONTINUOUS)
    return set_error(DFM_ERR_ARG, "jump conditions need the continuous adjoint");
  if (nseg > 0 && !lm_off) return set_error(DFM_ERR_ARG, "landmark offsets are NULL");
  Seg cs, ls;
  int rc = segs_of(nseg, cp_off, lm_off, n_, m_, cs, ls);
  if (rc) return rc;
  if ((rc = check_common(D, prec, sigma, cs.total, ls.total, T))) return rc;
  if (scheme != DFM_EULER && scheme != DFM_HEUN) return set_error(DFM_ERR_ARG, "unknown scheme %d", scheme);
  if (checkpoint_every < 0)
    return set_error(DFM_ERR_ARG, "checkpoint interval must be >= 1, or 0 for automatic");
  const long n = cs.total, m = ls.total;
  checkpoint_every = resolve_checkpoint(checkpoint_every, T, n, m, D);
  if (n == 0) return DFM_OK;
  if (!p0 || !a0 || (m > 0 && !x0)) return set_error(DFM_ERR_ARG, "NULL argument");
  if ((phase & 2) && (!grad_p0 || !grad_a0 || (m > 0 && !theta_T)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  const bool batch = nseg > 0;
  AdjointLayout need(nullptr, prec, D, cs, ls, T, checkpoint_every, batch);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  AdjointLayout L(work, prec, D, cs, ls, T, checkpoint_every, batch);
  Ctx cx{(hipStream_t)stream, D, sigma, prec, L.pw, L.pw_bytes, batch};
  if ((rc = upload_segs(cx, cs, ls, L.sa)) || (rc = clear_done(cx, L.c))) return rc;
  cx.done = L.c.done;
  const long nd = n * D, md = m * D, rec = 2 * nd + md;
  const double dt = T > 1 ? (tN - t0) / (T - 1) : 0.0;
  double *gate = L.sa.gate;
  // state record layout: [p (nd) | a (nd) | x (md)]
  auto P_ = [&](double *r) { return r; };
  auto A_ = [&](double *r) { return r + nd; };
  auto X_ = [&](double *r) { return r + 2 * nd; };

  if (phase & 1) {
    if ((rc = norm2(cx, a0, cs, gate))) return rc;
    // ---- forward: checkpoints at steps 0, k, 2k, ...
    double *cur = L.seg, *nxt = L.seg + rec;
    if ((rc = copy(cx, P_(cur), p0, nd)) || (rc = copy(cx, A_(cur), a0, nd))) return rc;
    if (m > 0 && (rc = copy(cx, X_(cur), x0, md))) return rc;
    for (int t = 0; t + 1 < T; t++) {
      if (t % L.k == 0 && (rc = copy(cx, L.ckpt + (size_t)(t / L.k) * rec, cur, rec))) return rc;
      if ((rc = shoot_step(cx, scheme, P_(cur), A_(cur), cs, dt, P_(nxt), A_(nxt), L.st, gate)))
        return rc;
      if (m > 0 && (rc = flow_step(cx, improved_euler, P_(cur), A_(cur), P_(nxt), A_(nxt), cs,
                                   X_(cur), ls, dt, X_(nxt), L.ft, gate)))
        return rc;
      std::swap(cur, nxt);
    }
    // the final state closes the last segment: the backward reads it instead of
    // recomputing it
    if ((rc = copy(cx, L.ckpt + (size_t)L.segs * rec, cur, rec))) return rc;
    if (x_final) *x_final = X_(cur);
  }
  if (!(phase & 2)) return DFM_OK;
  // terminal cotangents
  if ((rc = lin(cx, L.lp, p0, 0.0, nullptr, 0.0, nullptr, 0.0, nd))) return rc;
  if ((rc = lin(cx, L.la, a0, 0.0, nullptr, 0.0, nullptr, 0.0, nd))) return rc;
  if (m > 0) {
    if (!jumps)
      rc = copy(cx, L.th, theta_T, md);
    else if (jumps->terminal >= 0)
      rc = copy(cx, L.th, theta_T + (size_t)jumps->terminal * md, md);
    else
      rc = fill(cx.st, L.th, md, 0.0);
    if (rc) return rc;
  }

  // ---- backward, one checkpoint segment at a time: states tb and te come from the
  // checkpoints, the interior ones are recomputed (the same steps as the forward, so the
  // same values)
  for (int s = L.segs - 1; s >= 0; s--) {
    const int tb = s * L.k, te = std::min(T - 1, tb + L.k);
    double *S = L.seg;  // states tb..te
    if ((rc = copy(cx, S, L.ckpt + (size_t)s * rec, rec))) return rc;
    if ((rc = copy(cx, S + (size_t)(te - tb) * rec, L.ckpt + (size_t)(s + 1) * rec, rec)))
      return rc;
    for (int t = tb; t + 1 < te; t++) {
      double *a_ = S + (size_t)(t - tb) * rec, *b_ = S + (size_t)(t - tb + 1) * rec;
      if ((rc = shoot_step(cx, scheme, P_(a_), A_(a_), cs, dt, P_(b_), A_(b_), L.st, gate)))
        return rc;
      if (m > 0 && (rc = flow_step(cx, improved_euler, P_(a_), A_(a_), P_(b_), A_(b_), cs, X_(a_),
                                   ls, dt, X_(b_), L.ft, gate)))
        return rc;
    }
    for (int t = te - 1; t >= tb; t--) {
      double *zt = S + (size_t)(t - tb) * rec, *z1 = S + (size_t)(t - tb + 1) * rec;
      if (form == DFM_ADJOINT_CONTINUOUS) {
        const int j = jumps ? jumps->step_src[t + 1] : -1;
        const double *src = j >= 0 && m > 0 ? theta_T + (size_t)j * md : nullptr;
        if ((rc = continuous_step(cx, L, cs, ls, improved_euler, dt, P_(zt), A_(zt), X_(zt),
                                  P_(z1), A_(z1), X_(z1), src)))
          return rc;
        continue;
      }
      // flow step t: cotangent th of x_{t+1} -> thn of x_t, plus its pull onto z_t, z_{t+1}
      // (the elementwise updates ride on the sums' epilogues)
      if (m > 0) {
        if (improved_euler) {
          // xs = x_t + dt v(x_t; z_t)
          const Epi exs{X_(zt), 1.0, dt, nullptr, 0.0};
          if ((rc = conv_grad1(cx, X_(zt), nullptr, ls, P_(zt), A_(zt), cs, L.ft.xs, nullptr, &exs)))
            return rc;
          // ths = dt/2 J_x v(xs; z_{t+1})^T th
          const Epi ehs{nullptr, 0.0, 0.5 * dt, nullptr, 0.0};
          if ((rc = conv_grad1(cx, L.ft.xs, L.th, ls, P_(z1), A_(z1), cs, nullptr, L.ths, nullptr,
                               &ehs)))
            return rc;
          // lambda(z_{t+1}) += dt/2 J_z v(xs; z_{t+1})^T th
          const Epi ela{L.la, 1.0, 0.5 * dt, nullptr, 0.0}, elp{L.lp, 1.0, 0.5 * dt, nullptr, 0.0};
          if ((rc = conv_grad1(cx, P_(z1), A_(z1), cs, L.ft.xs, L.th, ls, L.la, L.lp, &ela, &elp)))
            return rc;
          // mu = dt/2 th + dt ths
          if ((rc = lin(cx, L.mu, L.th, 0.5 * dt, L.ths, dt, nullptr, 0.0, md))) return rc;
        } else {
          if ((rc = lin(cx, L.mu, L.th, dt, nullptr, 0.0, nullptr, 0.0, md))) return rc;
        }
        // thn = th + J_x v(x_t; z_t)^T mu (+ ths)
        const Epi etn{L.th, 1.0, 1.0, improved_euler ? L.ths : nullptr, 1.0};
        if ((rc = conv_grad1(cx, X_(zt), L.mu, ls, P_(zt), A_(zt), cs, nullptr, L.thn, nullptr,
                             &etn)))
          return rc;
        // flow pull onto z_t: J_z v(x_t; z_t)^T mu
        if ((rc = conv_grad1(cx, P_(zt), A_(zt), cs, X_(zt), L.mu, ls, L.laf, L.lpf))) return rc;
      }
      // shoot step t: cotangent (lp, la) of z_{t+1} -> (lpn, lan) of z_t, plus the flow pull
      if (scheme == DFM_EULER) {
        // lpn = lp + dt J_F(z_t)^T l (+ lpf), in the epilogue
        const Epi ep{L.lp, 1.0, dt, m > 0 ? L.lpf : nullptr, 1.0};
        const Epi ea{L.la, 1.0, dt, m > 0 ? L.laf : nullptr, 1.0};
        if ((rc = shoot_vjp(cx, P_(zt), A_(zt), L.lp, L.la, cs, L.lpn, L.lan, &ep, &ea))) return rc;
      } else {
        // z* = z_t + dt F(z_t)
        if ((rc = conv_grad1(cx, P_(zt), A_(zt), cs, P_(zt), A_(zt), cs, L.st.dpos, L.st.dmom))) return rc;
        if ((rc = lin(cx, L.st.ps, P_(zt), 1.0, L.st.dpos, dt, nullptr, 0.0, nd))) return rc;
        if ((rc = lin(cx, L.st.as, A_(zt), 1.0, L.st.dmom, -dt, nullptr, 0.0, nd))) return rc;
        // l* = dt/2 J_F(z*)^T l
        if ((rc = shoot_vjp(cx, L.st.ps, L.st.as, L.lp, L.la, cs, L.jsp, L.jsa))) return rc;
        // nu = dt/2 l + dt l* = dt/2 l + dt^2/2 J_F(z*)^T l
        if ((rc = lin(cx, L.nup, L.lp, 0.5 * dt, L.jsp, 0.5 * dt * dt, nullptr, 0.0, nd))) return rc;
        if ((rc = lin(cx, L.nua, L.la, 0.5 * dt, L.jsa, 0.5 * dt * dt, nullptr, 0.0, nd))) return rc;
        if ((rc = shoot_vjp(cx, P_(zt), A_(zt), L.nup, L.nua, cs, L.jp, L.ja))) return rc;
        // lpn = l + l* + J_F(z_t)^T nu (+ lpf)
        if ((rc = lin(cx, L.lpn, L.lp, 1.0, L.jsp, 0.5 * dt, L.jp, 1.0, nd))) return rc;
        if ((rc = lin(cx, L.lan, L.la, 1.0, L.jsa, 0.5 * dt, L.ja, 1.0, nd))) return rc;
        if (m > 0) {
          if ((rc = lin(cx, L.lpn, L.lpn, 1.0, L.lpf, 1.0, nullptr, 0.0, nd))) return rc;
          if ((rc = lin(cx, L.lan, L.lan, 1.0, L.laf, 1.0, nullptr, 0.0, nd))) return rc;
        }
      }
      std::swap(L.lp, L.lpn);
      std::swap(L.la, L.lan);
      if (m > 0) std::swap(L.th, L.thn);
    }
  }
  if ((rc = copy(cx, grad_p0, L.lp, nd)) || (rc = copy(cx, grad_a0, L.la, nd))) return rc;
  if (grad_x0 && m > 0 && (rc = copy(cx, grad_x0, L.th, md))) return rc;
  return DFM_OK;
}

// which: 0 shoot, 1 flow, 2 adjoint (checkpoint interval k)
size_t workspace_or_zero(int nseg, const long *cp_off, const long *lm_off, long n, long m, int D,
                         int prec, int T, int which, int k) {
  Seg cs, ls;
  if (segs_of(nseg, cp_off, lm_off, n, m, cs, ls)) return 0;
  if (check_common(D, prec, 1.0, cs.total, ls.total, T)) return 0;
  const bool batch = nseg > 0;
  if (which == 0) return ShootLayout(nullptr, prec, D, cs, batch).total;
  if (which == 1) return FlowLayout(nullptr, prec, D, cs, ls, batch).total;
  if (k < 0) return 0;
  k = resolve_checkpoint(k, T, cs.total, ls.total, D);
  return AdjointLayout(nullptr, prec, D, cs, ls, T, k, batch).total;
}

}  // namespace

int batch_adjoint_phase(hipStream_t st, int D, double sigma, int prec, int scheme, int improved,
                        int nseg, const long *cp_off, const long *lm_off, const double *p0,
                        const double *a0, const double *x0, int T, double t0, double tN, int k,
                        const double *theta_T, double *gp, double *ga, double *gx, void *work,
                        size_t work_bytes, int phase, const double **x_final, int form) {
  return adjoint_impl(st, D, sigma, prec, scheme, improved, nseg, cp_off, lm_off, p0, a0, 0, x0,
                      0, T, t0, tN, k, theta_T, gp, ga, gx, work, work_bytes, phase, x_final,
                      form);
}

size_t batch_adjoint_workspace(int D, int nseg, const long *cp_off, const long *lm_off, int T,
                               int k, int prec) {
  return workspace_or_zero(nseg, cp_off, lm_off, 0, 0, D, prec, T, 2, k);
}

}  // namespace dfm

using namespace dfm;

extern "C" {

size_t dfm_shoot_workspace(int dimpoint, long n, int precision) {
  return workspace_or_zero(0, nullptr, nullptr, n, 0, dimpoint, precision, 1, 0, 1);
}

int dfm_shoot(void *stream, int D, double sigma, int prec, int scheme, const double *p0,
              const double *a0, long n, int T, double t0, double tN, double *P, double *A,
              void *work, size_t work_bytes) {
  return shoot_impl(stream, D, sigma, prec, scheme, 0, nullptr, p0, a0, n, T, t0, tN, P, A, work,
                    work_bytes);
}

size_t dfm_flow_workspace(int dimpoint, long n, long m, int precision) {
  return workspace_or_zero(0, nullptr, nullptr, n, m, dimpoint, precision, 1, 1, 1);
}

int dfm_flow(void *stream, int D, double sigma, int prec, int improved_euler, const double *P,
             const double *A, long n, const double *x0, long m, int T, double t0, double tN,
             double *X, void *work, size_t work_bytes) {
  return flow_impl(stream, D, sigma, prec, improved_euler, 0, nullptr, nullptr, P, A, n, x0, m, T,
                   t0, tN, X, work, work_bytes);
}

size_t dfm_adjoint_workspace(int dimpoint, long n, long m, int T, int checkpoint_every,
                             int precision) {
  return workspace_or_zero(0, nullptr, nullptr, n, m, dimpoint, precision, T, 2, checkpoint_every);
}

int dfm_shoot_vjp(void *stream, int D, double sigma, int prec, const double *p, const double *a,
                  const double *mu_p, const double *mu_a, long n, double *jp, double *ja,
                  void *work, size_t work_bytes) {
  clear_error();
  int rc = check_common(D, prec, sigma, n, 0, 1);
  if (rc) return rc;
  if (n == 0) return DFM_OK;
  if (!p || !a || !mu_p || !mu_a || !jp || !ja) return set_error(DFM_ERR_ARG, "NULL argument");
  Ctx cx{(hipStream_t)stream, D, sigma, prec, work, work_bytes, false};
  return shoot_vjp(cx, p, a, mu_p, mu_a, Seg::single(n), jp, ja);
}

int dfm_shoot_flow_adjoint(void *stream, int D, double sigma, int prec, int scheme,
                           int improved_euler, const double *p0, const double *a0, long n,
                           const double *x0, long m, int T, double t0, double tN,
                           int checkpoint_every, const double *theta_T, double *grad_p0,
                           double *grad_a0, double *grad_x0, void *work, size_t work_bytes) {
  return adjoint_impl(stream, D, sigma, prec, scheme, improved_euler, 0, nullptr, nullptr, p0, a0,
                      n, x0, m, T, t0, tN, checkpoint_every, theta_T, grad_p0, grad_a0, grad_x0,
                      work, work_bytes);
}

int dfm_shoot_flow_adjoint_equations(void *stream, int D, double sigma, int prec, int scheme,
                                     int improved_euler, const double *p0, const double *a0,
                                     long n, const double *x0, long m, int T, double t0,
                                     double tN, int checkpoint_every, const double *theta_T,
                                     double *grad_p0, double *grad_a0, double *grad_x0,
                                     void *work, size_t work_bytes) {
  return adjoint_impl(stream, D, sigma, prec, scheme, improved_euler, 0, nullptr, nullptr, p0, a0,
                      n, x0, m, T, t0, tN, checkpoint_every, theta_T, grad_p0, grad_a0, grad_x0,
                      work, work_bytes, 3, nullptr, DFM_ADJOINT_CONTINUOUS);
}

int dfm_shoot_flow_adjoint_jumps(void *stream, int D, double sigma, int prec, int scheme,
                                 int improved_euler, const double *p0, const double *a0, long n,
                                 const double *x0, long m, int T, double t0, double tN,
                                 int checkpoint_every, int njumps, const int *jump_times,
                                 const double *conditions, double *grad_p0, double *grad_a0,
                                 double *grad_x0, void *work, size_t work_bytes) {
  clear_error();
  if (njumps < 1 || !jump_times)
    return set_error(DFM_ERR_ARG, "at least one jump time is needed");
  for (int j = 0; j < njumps; j++)
    if (jump_times[j] < 0 || jump_times[j] >= T)
      return set_error(DFM_ERR_ARG, "jump time %d outside [0, %d)", jump_times[j], T);
  const JumpPlan plan = plan_jumps(njumps, jump_times, T, improved_euler);
  return adjoint_impl(stream, D, sigma, prec, scheme, improved_euler, 0, nullptr, nullptr, p0, a0,
                      n, x0, m, T, t0, tN, checkpoint_every, conditions, grad_p0, grad_a0,
                      grad_x0, work, work_bytes, 3, nullptr, DFM_ADJOINT_CONTINUOUS, &plan);
}

// ---- batched subjects -----------------------------------------------------------------

size_t dfm_batch_shoot_workspace(int dimpoint, int nseg, const long *cp_offsets, int precision) {
  if (nseg < 1) return 0;
  return workspace_or_zero(nseg, cp_offsets, nullptr, 0, 0, dimpoint, precision, 1, 0, 1);
}

int dfm_batch_shoot(void *stream, int D, double sigma, int prec, int scheme, int nseg,
                    const long *cp_offsets, const double *p0, const double *a0, int T, double t0,
                    double tN, double *P, double *A, void *work, size_t work_bytes) {
  if (nseg < 1) return set_error(DFM_ERR_ARG, "segment count must be >= 1");
  return shoot_impl(stream, D, sigma, prec, scheme, nseg, cp_offsets, p0, a0, 0, T, t0, tN, P, A,
                    work, work_bytes);
}

size_t dfm_batch_flow_workspace(int dimpoint, int nseg, const long *cp_offsets,
                                const long *lm_offsets, int precision) {
  if (nseg < 1 || !lm_offsets) return 0;
  return workspace_or_zero(nseg, cp_offsets, lm_offsets, 0, 0, dimpoint, precision, 1, 1, 1);
}

int dfm_batch_flow(void *stream, int D, double sigma, int prec, int improved_euler, int nseg,
                   const long *cp_offsets, const long *lm_offsets, const double *P,
                   const double *A, const double *x0, int T, double t0, double tN, double *X,
                   void *work, size_t work_bytes) {
  if (nseg < 1) return set_error(DFM_ERR_ARG, "segment count must be >= 1");
  return flow_impl(stream, D, sigma, prec, improved_euler, nseg, cp_offsets, lm_offsets, P, A, 0,
                   x0, 0, T, t0, tN, X, work, work_bytes);
}

size_t dfm_batch_adjoint_workspace(int dimpoint, int nseg, const long *cp_offsets,
                                   const long *lm_offsets, int T, int checkpoint_every,
                                   int precision) {
  if (nseg < 1 || !lm_offsets) return 0;
  return workspace_or_zero(nseg, cp_offsets, lm_offsets, 0, 0, dimpoint, precision, T, 2,
                           checkpoint_every);
}

int dfm_batch_shoot_flow_adjoint(void *stream, int D, double sigma, int prec, int scheme,
                                 int improved_euler, int nseg, const long *cp_offsets,
                                 const long *lm_offsets, const double *p0, const double *a0,
                                 const double *x0, int T, double t0, double tN,
                                 int checkpoint_every, const double *theta_T, double *grad_p0,
                                 double *grad_a0, double *grad_x0, void *work,
                                 size_t work_bytes) {
  if (nseg < 1) return set_error(DFM_ERR_ARG, "segment count must be >= 1");
  return adjoint_impl(stream, D, sigma, prec, scheme, improved_euler, nseg, cp_offsets,
                      lm_offsets, p0, a0, 0, x0, 0, T, t0, tN, checkpoint_every, theta_T, grad_p0,
                      grad_a0, grad_x0, work, work_bytes);
}

}  // extern "C"
