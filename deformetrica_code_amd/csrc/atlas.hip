// atlas.hip — the deterministic atlas log-likelihood and its gradient for a population of
// subjects, every subject in the same launches: triangle surfaces in 3D, or polylines
// (segments) in 2D or 3D.
//
// Model (DeterministicAtlas.cxx:100-151, 159-281; AbstractAtlas.cxx:490-620, 727-751):
//   subject s: control points cp (shared), momenta a_s, template T0 flowed by the geodesic
//   shot from (cp, a_s) to T_s = phi_s(T0);
//   loglik[0] = -1/2 sum_s r_s / sigma_data^2,  r_s = |T_s - target_s|^2 (current or
//               varifold norm at the attachment width)
//   loglik[1] = -1/2 sum_s a_s^T K(cp, cp) a_s  (RKHS regularity)
//   gradients of loglik[0] + loglik[1] with respect to a_s, cp and the template vertices:
//     d/d a_s  = -adjoint_a_s - K a_s
//     d/d cp   = -sum_s (adjoint_p_s + Grad1(cp, a_s))
//     d/d T0   = -sum_s adjoint_x_s, smoothed by Convolve at the smoothing width when
//               that width is > 1e-20 (AbstractAtlas::ConvolveGradTemplate)
//   where adjoint_* is the gradient of r_s / (2 sigma_data^2) through the shot and the flow:
//   the reference's continuous adjoint equations (AdjointEquationsIntegrator) with
//   desc.adjoint = DFM_ADJOINT_CONTINUOUS, or the exact discrete adjoint of the forward
//   scheme (equal to O(dt)) with DFM_ADJOINT_DISCRETE.
//
// One gradient call: batched shoot + flow forward with checkpoints; data-term geometry of
// all deformed templates; two fused Convolve + ConvolveGradient pair sums over all subjects
// (template-template and template-target, targets ragged); per-cell corner contributions
// and a deterministic per-vertex gather; the batched adjoint backward; the regularity pair
// sum; deterministic reductions over subjects.  Everything is allocated at creation, so a
// gradient call launches kernels only and can be captured in a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "dfm_hip.h"
#include "dfm_internal.h"

namespace dfm {
namespace {

size_t a256(size_t v) { return (v + 255) & ~(size_t)255; }
unsigned blocks(long n) { return (unsigned)std::max(1L, (n + 255) / 256); }

// dst[s * len + k] = src[k]
__global__ __launch_bounds__(256) void k_tile(double *dst, const double *src, long len, int nsub) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len * nsub) return;
  dst[i] = src[i % len];
}

// dst[s * m * D + i * D + k] = src[order[i] * D + k]: the caller's template rows in the
// atlas's vertex order, once per subject
__global__ __launch_bounds__(256) void k_tile_rows(double *dst, const double *src,
                                                   const int *order, long m, int D, int nsub) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m * D * nsub) return;
  const long r = (i / D) % m;
  dst[i] = src[(long)order[r] * D + i % D];
}

// dst[order[i] * D + k] = src[i * D + k]: rows in the atlas's vertex order back to the
// caller's
__global__ __launch_bounds__(256) void k_scatter_rows(double *dst, const double *src,
                                                      const int *order, long m, int D) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m * D) return;
  dst[(long)order[i / D] * D + i % D] = src[i];
}

// out[k] = c * sum_s (a[s * len + k] + b[s * len + k])  (b may be null), subjects in order
__global__ __launch_bounds__(256) void k_sum_subjects(double *out, const double *a,
                                                      const double *b, long len, int nsub,
                                                      double c) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len) return;
  double v = 0.0;
  for (int s = 0; s < nsub; s++) {
    v += a[s * len + k];
    if (b) v += b[s * len + k];
  }
  out[k] = c * v;
}

// out = -(a + b), elementwise
__global__ __launch_bounds__(256) void k_neg_sum(double *out, const double *a, const double *b,
                                                 long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = -(a[i] + b[i]);
}

// residuals r_s = T2 + S2 - 2 ST;  loglik = (-1/2 sum r_s / sigma2, -1/2 sum reg_s)
__global__ __launch_bounds__(64) void k_loglik(const double *T2, const double *S2,
                                               const double *ST, const double *reg, int nsub,
                                               double inv_sigma2, double *residuals,
                                               double *loglik) {
  if (threadIdx.x != 0) return;
  double data = 0.0, r = 0.0;
  for (int s = 0; s < nsub; s++) {
    const double res = T2[s] + S2[s] - 2.0 * ST[s];
    if (residuals) residuals[s] = res;
    data += res;
    r += reg[s];
  }
  loglik[0] = -0.5 * data * inv_sigma2;
  loglik[1] = -0.5 * r;
}

int launched(const char *what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

KParams blank() {
  KParams p;
  std::memset(&p, 0, sizeof(p));
  return p;
}

template <typename T>
int dmalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) return DFM_OK;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(DFM_ERR_NOMEM, "hipMalloc(%zu): %s", count * sizeof(T), hipGetErrorString(e));
  }
  return DFM_OK;
}

template <typename T>
int upload(T **p, const T *h, size_t count) {
  int rc = dmalloc(p, count);
  if (rc || count == 0) return rc;
  hipError_t e = hipMemcpy(*p, h, count * sizeof(T), hipMemcpyHostToDevice);
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "H2D: %s", hipGetErrorString(e));
}

}  // namespace
}  // namespace dfm

using namespace dfm;

struct dfm_atlas {
  dfm_atlas_desc c{};
  int dv = 3;
  int D = 3, corners = 3;  // dimension, points per cell (3 triangles, 2 segments)
  long F = 0, m = 0, n = 0;
  int S = 0;
  std::vector<long> cp_off, lm_off, face_off, tgt_off;
  Seg face_seg, tgt_seg, cp_seg;
  // device data fixed at creation
  int *faces = nullptr, *inc_off = nullptr, *inc = nullptr;
  // internal template vertex i is the caller's vertex order[i] (nullptr: the caller's order)
  int *order = nullptr;
  long *d_face_off = nullptr, *d_tgt_off = nullptr, *d_cp_off = nullptr;
  double *cT = nullptr, *wT = nullptr, *aT = nullptr, *T2 = nullptr;
  // gradient-call arena
  char *work = nullptr;
  size_t work_bytes = 0;
  std::vector<void *> owned;

  struct Layout {
    double *p0r, *x0r, *cS, *wS, *aS, *nS, *k1, *k2, *g1, *g2, *contrib, *theta;
    double *gp, *ga, *gx, *ka, *gr, *S2, *ST, *reg, *bad, *gl2;
    void *adj, *pw;
    size_t adj_bytes, pw_bytes, total;
  };

  Layout layout(char *base) const {
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t count) {
      double *p = base ? (double *)(base + off) : nullptr;
      off += a256(count * sizeof(double));
      return p;
    };
    const long SF = (long)S * F, Sn = (long)S * n, Sm = (long)S * m;
    L.p0r = take(Sn * D); L.x0r = take(Sm * D);
    L.cS = take(SF * D); L.wS = take(SF * dv); L.aS = take(SF * dv); L.nS = take(SF * D);
    L.k1 = take(SF * dv); L.k2 = take(SF * dv); L.g1 = take(SF * D); L.g2 = take(SF * D);
    L.contrib = take(SF * corners * D); L.theta = take(Sm * D);
    L.gp = take(Sn * D); L.ga = take(Sn * D); L.gx = take(Sm * D);
    L.ka = take(Sn * D); L.gr = take(Sn * D);
    L.S2 = take(S); L.ST = take(S); L.reg = take(S); L.bad = take(1); L.gl2 = take(m * D);
    L.adj_bytes = batch_adjoint_workspace(D, S, cp_off.data(), lm_off.data(), c.number_of_time_points,
                                          c.checkpoint_every, c.precision);
    L.adj = base ? (void *)(base + off) : nullptr;
    off += a256(L.adj_bytes);
    size_t pw = 0;
    pw = std::max(pw, pairsum_batch_workspace(data_grad_op(c.kind), c.precision, D, dv, face_seg,
                                              face_seg));
    pw = std::max(pw, pairsum_batch_workspace(data_grad_op(c.kind), c.precision, D, dv, face_seg,
                                              tgt_seg));
    pw = std::max(pw, pairsum_batch_workspace(OP_CONVGRAD1, c.precision, D, D, cp_seg, cp_seg));
    pw = std::max(pw, pairsum_batch_workspace(OP_CONV, c.precision, D, D, Seg::single(m),
                                              Seg::single(m)));
    L.pw_bytes = pw;
    L.pw = base ? (void *)(base + off) : nullptr;
    off += a256(pw);
    L.total = off;
    return L;
  }

  ~dfm_atlas() {
    for (void *p : owned)
      if (p) (void)hipFree(p);
  }
};

namespace {

int check_desc(const dfm_atlas_desc *d) {
  if (!d) return set_error(DFM_ERR_ARG, "descriptor is NULL");
  if (d->kind != DFM_CURRENT && d->kind != DFM_VARIFOLD)
    return set_error(DFM_ERR_ARG, "unknown data term %d", d->kind);
  if (!valid_precision(d->precision))
    return set_error(DFM_ERR_ARG, "unknown precision %d", d->precision);
  if (d->scheme != DFM_EULER && d->scheme != DFM_HEUN)
    return set_error(DFM_ERR_ARG, "unknown scheme %d", d->scheme);
  if (d->number_of_time_points < 1 || d->checkpoint_every < 0)
    return set_error(DFM_ERR_ARG, "time points must be >= 1 and the checkpoint interval "
                                  ">= 1, or 0 for automatic");
  if (d->adjoint != DFM_ADJOINT_DISCRETE && d->adjoint != DFM_ADJOINT_CONTINUOUS)
    return set_error(DFM_ERR_ARG, "unknown adjoint form %d", d->adjoint);
  if (!(d->deformation_width > 0) || !(d->attachment_width > 0) || !(d->data_sigma_squared > 0))
    return set_error(DFM_ERR_ARG, "kernel widths and the data variance must be positive");
  if (d->nsubjects < 1 || d->nsubjects > 65535)
    return set_error(DFM_ERR_ARG, "subject count must be in [1, 65535]");
  if (d->n_control_points < 1 || d->n_template_vertices < 1 || d->n_template_faces < 1)
    return set_error(DFM_ERR_ARG, "empty control points or template");
  if (!d->template_faces || !d->target_vertex_offsets || !d->target_face_offsets ||
      !d->target_vertices || !d->target_faces)
    return set_error(DFM_ERR_ARG, "NULL mesh argument");
  const int D = d->dimension ? d->dimension : 3, corners = d->cell_corners ? d->cell_corners : 3;
  if (!((corners == 3 && D == 3) || (corners == 2 && (D == 2 || D == 3))))
    return set_error(DFM_ERR_DIM, "cells of %d points in dimension %d: triangles need 3, "
                                  "segments 2 or 3", corners, D);
  return DFM_OK;
}

}  // namespace

extern "C" {

int dfm_atlas_create(const dfm_atlas_desc *desc, dfm_atlas **out) {
  clear_error();
  if (!out) return set_error(DFM_ERR_ARG, "out is NULL");
  *out = nullptr;
  int rc = check_desc(desc);
  if (rc) return rc;
  std::unique_ptr<dfm_atlas> A(new dfm_atlas());
  A->c = *desc;
  A->D = desc->dimension ? desc->dimension : 3;
  A->corners = desc->cell_corners ? desc->cell_corners : 3;
  const int D = A->D, K = A->corners;
  // weights per cell: the normal or tangent (current), the upper triangle of its square
  // (varifold)
  A->dv = D;  // normals (current) or unit-root normals u (varifold, OP_VARIFOLD)
  A->S = desc->nsubjects;
  A->n = desc->n_control_points;
  A->m = desc->n_template_vertices;
  A->F = desc->n_template_faces;
  const int S = A->S;
  for (int s = 0; s <= S; s++) {
    A->cp_off.push_back((long)s * A->n);
    A->lm_off.push_back((long)s * A->m);
    A->face_off.push_back((long)s * A->F);
  }
  // template topology (this checks every index), then the template vertices in the order
  // of their first use by the cells: a mesh lists its cells in a spatially coherent order
  // more often than its vertices (a subdivided icosphere appends each level's midpoints
  // after every coarser vertex), and the flow's sums take the template vertices as
  // targets, whose wavefronts are then compact.  On the 64-subject atlas the FP32C gradient
  // takes 74.9 ms in this order against 77.9 ms in the icosphere's own.
  std::vector<int> off(A->m + 1), inc((size_t)K * A->F);
  if ((rc = dfm_cell_incidence(desc->template_faces, A->F, K, A->m, off.data(), inc.data())))
    return rc;
  std::vector<int> order, rank(A->m, -1), cells(desc->template_faces,
                                                desc->template_faces + (size_t)K * A->F);
  order.reserve(A->m);
  for (int &v : cells) {
    if (rank[v] < 0) {
      rank[v] = (int)order.size();
      order.push_back(v);
    }
    v = rank[v];
  }
  for (long v = 0; v < A->m; v++)
    if (rank[v] < 0) {
      rank[v] = (int)order.size();
      order.push_back((int)v);
    }
  bool same = true;
  for (long i = 0; i < A->m && same; i++) same = order[i] == i;
  if (!same && (rc = dfm_cell_incidence(cells.data(), A->F, K, A->m, off.data(), inc.data())))
    return rc;
  // targets: local face indices -> indices into the concatenated vertices
  const long *vo = desc->target_vertex_offsets, *fo = desc->target_face_offsets;
  if (vo[0] != 0 || fo[0] != 0) return set_error(DFM_ERR_ARG, "target offsets must start at 0");
  std::vector<int> tf((size_t)K * fo[S]);
  for (int s = 0; s < S; s++) {
    if (vo[s + 1] < vo[s] || fo[s + 1] < fo[s])
      return set_error(DFM_ERR_ARG, "target offsets must be non-decreasing");
    const long nvs = vo[s + 1] - vo[s];
    for (long i = K * fo[s]; i < K * fo[s + 1]; i++) {
      const int v = desc->target_faces[i];
      if (v < 0 || v >= nvs)
        return set_error(DFM_ERR_ARG, "target %d: cell %ld references vertex %d outside [0, %ld)", s,
                         i / K - fo[s], v, nvs);
      tf[i] = (int)(v + vo[s]);
    }
  }
  if (vo[S] > 0x7fffffffL || fo[S] > 0x7fffffffL)
    return set_error(DFM_ERR_ARG, "target counts must be below 2^31");
  A->tgt_off.assign(fo, fo + S + 1);
  // device copies
  auto own = [&](void *p) { A->owned.push_back(p); };
  if ((rc = upload(&A->faces, cells.data(), (size_t)K * A->F))) return rc;
  own(A->faces);
  if (!same) {
    if ((rc = upload(&A->order, order.data(), order.size()))) return rc;
    own(A->order);
  }
  if ((rc = upload(&A->inc_off, off.data(), off.size()))) return rc;
  own(A->inc_off);
  if ((rc = upload(&A->inc, inc.data(), inc.size()))) return rc;
  own(A->inc);
  if ((rc = upload(&A->d_face_off, A->face_off.data(), S + 1))) return rc;
  own(A->d_face_off);
  if ((rc = upload(&A->d_tgt_off, A->tgt_off.data(), S + 1))) return rc;
  own(A->d_tgt_off);
  if ((rc = upload(&A->d_cp_off, A->cp_off.data(), S + 1))) return rc;
  own(A->d_cp_off);
  if ((rc = seg_from_host(A->face_seg, S, A->face_off.data(), A->d_face_off))) return rc;
  if ((rc = seg_from_host(A->tgt_seg, S, A->tgt_off.data(), A->d_tgt_off))) return rc;
  if ((rc = seg_from_host(A->cp_seg, S, A->cp_off.data(), A->d_cp_off))) return rc;
  // target geometry and squared norms, once
  const long ntf = fo[S];
  const int dv = A->dv;
  if ((rc = dmalloc(&A->cT, ntf * D)) || (own(A->cT), rc = dmalloc(&A->wT, ntf * dv)) ||
      (own(A->wT), rc = dmalloc(&A->aT, ntf * dv)) || (own(A->aT), rc = dmalloc(&A->T2, S)))
    return rc;
  own(A->T2);
  double *tv = nullptr, *nT = nullptr, *kT = nullptr, *bad = nullptr;
  int *tfd = nullptr;
  void *pw = nullptr;
  const size_t pwb =
      pairsum_batch_workspace(data_conv_op(desc->kind), desc->precision, D, dv, A->tgt_seg,
                              A->tgt_seg);
  struct Tmp {
    std::vector<void *> v;
    ~Tmp() {
      for (void *p : v)
        if (p) (void)hipFree(p);
    }
  } tmp;
  if ((rc = upload(&tv, desc->target_vertices, (size_t)D * vo[S]))) return rc;
  tmp.v.push_back(tv);
  if ((rc = upload(&tfd, tf.data(), tf.size()))) return rc;
  tmp.v.push_back(tfd);
  if ((rc = dmalloc(&nT, ntf * D)) || (tmp.v.push_back(nT), rc = dmalloc(&kT, ntf * dv)) ||
      (tmp.v.push_back(kT), rc = dmalloc(&bad, 1)))
    return rc;
  tmp.v.push_back(bad);
  if (pwb) {
    if ((rc = dmalloc((char **)&pw, pwb))) return rc;
    tmp.v.push_back(pw);
  }
  hipStream_t st = nullptr;
  if ((rc = fill(st, bad, 1, 0.0))) return rc;
  if ((rc = cells_geometry(st, desc->kind, D, K, tv, vo[S], tfd, ntf, 1, A->cT, A->wT, A->aT, nT,
                           bad)))
    return rc;
  KParams p = blank();
  p.x = A->cT; p.ta = A->aT; p.y = A->cT; p.w = A->wT; p.out0 = kT;
  if ((rc = pairsum_batch(st, data_conv_op(desc->kind), desc->precision, D, dv,
                          desc->attachment_width, p, A->tgt_seg, A->tgt_seg, ntf * dv, 0, pw, pwb)))
    return rc;
  if ((rc = seg_dot(st, kT, A->aT, dv, S, 0, A->d_tgt_off, A->T2))) return rc;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "atlas setup: %s", hipGetErrorString(e));
  // gradient-call arena
  dfm_atlas::Layout need = A->layout(nullptr);
  if ((rc = dmalloc(&A->work, need.total))) return rc;
  own(A->work);
  A->work_bytes = need.total;
  *out = A.release();
  return DFM_OK;
}

int dfm_atlas_destroy(dfm_atlas *atlas) {
  clear_error();
  delete atlas;
  return DFM_OK;
}

size_t dfm_atlas_device_bytes(const dfm_atlas *atlas) { return atlas ? atlas->work_bytes : 0; }

int dfm_atlas_gradient(dfm_atlas *A, void *stream, const double *cp, const double *momenta,
                       const double *template_vertices, double *loglik, double *residuals,
                       double *grad_cp, double *grad_momenta, double *grad_template) {
  clear_error();
  if (!A) return set_error(DFM_ERR_ARG, "atlas is NULL");
  if (!cp || !momenta || !template_vertices || !loglik)
    return set_error(DFM_ERR_ARG, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  const dfm_atlas_desc &c = A->c;
  const int S = A->S, dv = A->dv, prec = c.precision, D = A->D, K = A->corners;
  const long n = A->n, m = A->m, F = A->F, SF = (long)S * F;
  dfm_atlas::Layout L = A->layout(A->work);
  int rc;
  // per-subject copies of the shared control points and template
  hipLaunchKernelGGL(k_tile, dim3(blocks(n * D * S)), dim3(256), 0, st, L.p0r, cp, n * D, S);
  if (A->order)
    hipLaunchKernelGGL(k_tile_rows, dim3(blocks(m * D * S)), dim3(256), 0, st, L.x0r,
                       template_vertices, A->order, m, D, S);
  else
    hipLaunchKernelGGL(k_tile, dim3(blocks(m * D * S)), dim3(256), 0, st, L.x0r,
                       template_vertices, m * D, S);
  if ((rc = launched("k_tile"))) return rc;
  // forward: shoot + flow of every subject, checkpointed
  const double *xT = nullptr;
  if ((rc = batch_adjoint_phase(st, D, c.deformation_width, prec, c.scheme, c.improved_euler, S,
                                A->cp_off.data(), A->lm_off.data(), L.p0r, momenta, L.x0r,
                                c.number_of_time_points, c.t0, c.tN, c.checkpoint_every, nullptr,
                                nullptr, nullptr, nullptr, L.adj, L.adj_bytes, 1, &xT, c.adjoint)))
    return rc;
  // data term of every deformed template against its target
  if ((rc = fill(st, L.bad, 1, 0.0))) return rc;
  if ((rc = cells_geometry(st, c.kind, D, K, xT, m, A->faces, F, S, L.cS, L.wS, L.aS, L.nS,
                           L.bad)))
    return rc;
  KParams p = blank();
  p.x = L.cS; p.ta = L.aS; p.y = L.cS; p.w = L.wS; p.out0 = L.k1; p.out1 = L.g1;
  if ((rc = pairsum_batch(st, data_grad_op(c.kind), prec, D, dv, c.attachment_width, p,
                          A->face_seg, A->face_seg, SF * dv, SF * D, L.pw, L.pw_bytes)))
    return rc;
  p.y = A->cT; p.w = A->wT; p.out0 = L.k2; p.out1 = L.g2;
  if ((rc = pairsum_batch(st, data_grad_op(c.kind), prec, D, dv, c.attachment_width, p,
                          A->face_seg, A->tgt_seg, SF * dv, SF * D, L.pw, L.pw_bytes)))
    return rc;
  if ((rc = seg_dot(st, L.k1, L.aS, dv, S, F, nullptr, L.S2))) return rc;
  if ((rc = seg_dot(st, L.k2, L.aS, dv, S, F, nullptr, L.ST))) return rc;
  // regularity: K a_s and Grad1(cp, a_s) in one pair sum over all subjects
  p = blank();
  p.x = L.p0r; p.ta = momenta; p.y = L.p0r; p.w = momenta; p.out0 = L.ka; p.out1 = L.gr;
  if ((rc = pairsum_batch(st, OP_CONVGRAD1, prec, D, D, c.deformation_width, p, A->cp_seg,
                          A->cp_seg, (long)S * n * D, (long)S * n * D, L.pw, L.pw_bytes)))
    return rc;
  if ((rc = seg_dot(st, L.ka, momenta, D, S, n, nullptr, L.reg))) return rc;
  hipLaunchKernelGGL(k_loglik, dim3(1), dim3(64), 0, st, A->T2, L.S2, L.ST, L.reg, S,
                     1.0 / c.data_sigma_squared, residuals, loglik);
  if ((rc = launched("k_loglik"))) return rc;
  if (!grad_cp && !grad_momenta && !grad_template) return DFM_OK;
  if ((rc = cells_contrib(st, c.kind, D, K, xT, m, A->faces, F, S, L.nS, L.k1, L.g1, L.k2, L.g2,
                          L.contrib)))
    return rc;
  // theta = d (r_s / (2 sigma_data^2)) / d x_T
  if ((rc = cells_gather(st, D, K, A->inc_off, A->inc, m, F, S, L.contrib,
                         0.5 / c.data_sigma_squared, L.theta)))
    return rc;
  // backward through the flow and the shot
  if ((rc = batch_adjoint_phase(st, D, c.deformation_width, prec, c.scheme, c.improved_euler, S,
                                A->cp_off.data(), A->lm_off.data(), L.p0r, momenta, L.x0r,
                                c.number_of_time_points, c.t0, c.tN, c.checkpoint_every, L.theta,
                                L.gp, L.ga, L.gx, L.adj, L.adj_bytes, 2, nullptr, c.adjoint)))
    return rc;
  if (grad_momenta)
    hipLaunchKernelGGL(k_neg_sum, dim3(blocks((long)S * n * D)), dim3(256), 0, st, grad_momenta,
                       L.ga, L.ka, (long)S * n * D);
  if (grad_cp)
    hipLaunchKernelGGL(k_sum_subjects, dim3(blocks(n * D)), dim3(256), 0, st, grad_cp, L.gp, L.gr,
                       n * D, S, -1.0);
  if ((rc = launched("atlas gradient assembly"))) return rc;
  if (grad_template) {
    // in the atlas's vertex order (L.x0r holds the template in it) until the last step
    const bool smooth = c.smoothing_width > 1e-20;
    double *g = (smooth || A->order) ? L.gl2 : grad_template;
    hipLaunchKernelGGL(k_sum_subjects, dim3(blocks(m * D)), dim3(256), 0, st, g, L.gx,
                       (const double *)nullptr, m * D, S, -1.0);
    if ((rc = launched("k_sum_subjects"))) return rc;
    if (smooth) {
      p = blank();
      p.x = L.x0r; p.y = L.x0r; p.w = L.gl2;
      p.out0 = A->order ? L.gx : grad_template;  // L.gx is free once summed
      if ((rc = pairsum_batch(st, OP_CONV, prec, D, D, c.smoothing_width, p, Seg::single(m),
                              Seg::single(m), m * D, 0, L.pw, L.pw_bytes)))
        return rc;
      g = L.gx;
    }
    if (A->order) {
      hipLaunchKernelGGL(k_scatter_rows, dim3(blocks(m * D)), dim3(256), 0, st, grad_template,
                         g, A->order, m, D);
      if ((rc = launched("k_scatter_rows"))) return rc;
    }
  }
  return DFM_OK;
}

}  // extern "C"
