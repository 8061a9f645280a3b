// surfaces.hip — triangle-surface, polyline, volume-mesh, point-cloud and landmark data terms
// on the GPU: currents (OrientedSurfaceMesh, OrientedPolyLine), varifolds
// (NonOrientedSurfaceMesh, NonOrientedPolyLine), volume measures (OrientedVolumeMesh),
// weighted Diracs (PointCloud) and squared distances (Landmark), their match values and
// their gradients with respect to the deformed template's vertices.  The polyline path (segments in 2D or 3D) follows the same plan with
// k_seg_geometry / k_seg_contrib in place of the face kernels.
//
// Every kernel sum is one pair-sum call; the rest is per-face and per-vertex work:
//   k_face_geometry   centers, kernel weights W_f and per-face coefficients A_f
//                     current:  W = A = N = (p2 - p0) x (p1 - p0) / 2
//                               (OrientedSurfaceMesh.cxx:298-326)
//                     varifold: W = A = N = u, the normal divided by sqrt(|N| + 1e-20)
//                               (NonOrientedSurfaceMesh.cxx:269-308)
//   match             |S|^2 + |T|^2 - 2 <S, T> with <X, Y> = sum_f A^Y_f . (K W^X)(c^Y_f):
//                     three Convolve calls and deterministic two-stage FP64 reductions
//                     (OrientedSurfaceMesh.cxx:125-163, NonOrientedSurfaceMesh.cxx:124-165);
//                     for the varifold the Convolve is the first half of OP_VARIFOLD,
//                     V_f = sum_j K (u_f . u_j) u_j, and u_f . V_f = sum_j K (u_f . u_j)^2 is
//                     the reference's special_product of N N^T with (K W)(c_f) for its
//                     W_j = upper triangle of u_j u_j^T (:364-379)
//   gradient          (K W)(c_S) and ConvolveGradient(c_S, A) from the template and from
//                     the target (two fused calls: Convolve + Grad1 for the current,
//                     OP_VARIFOLD for the varifold, whose Grad1 part is the reference's
//                     contraction of the varlin list with N N^T (:243-249) and whose
//                     Convolve part is S((K W)(c_f)) u_f, all the contraction needs),
//                     then per-face corner contributions (k_face_contrib) and a
//                     deterministic per-vertex gather over a vertex->corner incidence list
//                     (k_vertex_gather), in place of the reference's serial scatter
//                     (OrientedSurfaceMesh.cxx:205-234, NonOrientedSurfaceMesh.cxx:213-256).
// Faces with an index outside [0, nv) contribute nothing and set a flag that the
// entry points report (device-side bounds checks: no out-of-range access).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dfm_hip.h"
#include "dfm_internal.h"

namespace dfm {
namespace {

constexpr int kRedBlocks = 256;

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Geometry of each face; bad indices flag the face (weights zero).  nsub meshes share
// one triangle list F (nf faces over nv vertices): face f of the output belongs to mesh
// f / nf, whose vertices start at row (f / nf) * nv of V.
__global__ __launch_bounds__(256) void k_face_geometry(int kind, const double *V, long nv,
                                                       const int *F, long nf, int nsub,
                                                       double *C, double *W, double *A,
                                                       double *Nout, double *bad) {
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf * nsub) return;
  const long sub = f / nf, fl = f - sub * nf;
  V += sub * nv * 3;
  int i0 = F[3 * fl], i1 = F[3 * fl + 1], i2 = F[3 * fl + 2];
  if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= nv || i1 >= nv || i2 >= nv) {
    for (int k = 0; k < 3; k++) C[3 * f + k] = 0.0, Nout[3 * f + k] = 0.0;
    for (int k = 0; k < 3; k++) W[3 * f + k] = 0.0, A[3 * f + k] = 0.0;
    *bad = 1.0;
    return;
  }
  double p0[3], p1[3], p2[3], u[3], v[3], n[3];
  for (int k = 0; k < 3; k++) {
    p0[k] = V[3 * (long)i0 + k];
    p1[k] = V[3 * (long)i1 + k];
    p2[k] = V[3 * (long)i2 + k];
    C[3 * f + k] = (p0[k] + p1[k] + p2[k]) / 3.0;
    u[k] = p2[k] - p0[k];
    v[k] = p1[k] - p0[k];
  }
  cross3(u, v, n);
  for (int k = 0; k < 3; k++) n[k] /= 2;
  if (kind == DFM_CURRENT) {
    for (int k = 0; k < 3; k++) W[3 * f + k] = A[3 * f + k] = Nout[3 * f + k] = n[k];
    return;
  }
  const double s = sqrt(sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) + 1e-20);
  for (int k = 0; k < 3; k++) W[3 * f + k] = A[3 * f + k] = Nout[3 * f + k] = n[k] / s;
}

// partial[b] = sum over the block's faces of dot(X_f, Y_f) (dv columns)
__global__ __launch_bounds__(256) void k_dot_partial(const double *X, const double *Y, long nf,
                                                     int dv, double *partial) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long f = (long)blockIdx.x * blockDim.x + threadIdx.x; f < nf;
       f += (long)gridDim.x * blockDim.x)
    for (int k = 0; k < dv; k++) s += X[dv * f + k] * Y[dv * f + k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out = c0 * sum(p0) + c1 * sum(p1) + c2 * sum(p2) over kRedBlocks partials each
__global__ __launch_bounds__(256) void k_combine(const double *p0, double c0, const double *p1,
                                                 double c1, const double *p2, double c2,
                                                 double *out) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  sh[t] = c0 * p0[t] + c1 * p1[t] + c2 * p2[t];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) sh[t] += sh[t + st];
    __syncthreads();
  }
  if (t == 0) *out = sh[0];
}

// contrib[f][c] (3 corners x 3) of the match gradient (nsub meshes as in k_face_geometry)
__global__ __launch_bounds__(256) void k_face_contrib(int kind, const double *V, long nv,
                                                      const int *F, long nf, int nsub,
                                                      const double *N, const double *ktS,
                                                      const double *gS, const double *ktT,
                                                      const double *gT, double *contrib) {
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf * nsub) return;
  const long sub = f / nf, fl = f - sub * nf;
  V += sub * nv * 3;
  double *out = contrib + 9 * f;
  for (int k = 0; k < 9; k++) out[k] = 0.0;
  const int i0 = F[3 * fl], i1 = F[3 * fl + 1], i2 = F[3 * fl + 2];
  if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= nv || i1 >= nv || i2 >= nv) return;
  double e[3][3], kt[3], gk[3];
  for (int k = 0; k < 3; k++) {
    e[0][k] = V[3 * (long)i2 + k] - V[3 * (long)i1 + k];
    e[1][k] = V[3 * (long)i0 + k] - V[3 * (long)i2 + k];
    e[2][k] = V[3 * (long)i1 + k] - V[3 * (long)i0 + k];
    gk[k] = (gS[3 * f + k] - gT[3 * f + k]) * 2.0 / 3.0;
  }
  if (kind == DFM_CURRENT) {
    for (int k = 0; k < 3; k++) kt[k] = ktS[3 * f + k] - ktT[3 * f + k];
  } else {
    const double *n = N + 3 * f;
    const double n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (!(n2 > 1e-20)) return;
    // t = 2 S((K W_S)(c_f) - (K W_T)(c_f)) N = 2 (V_S - V_T), the Convolve parts of the
    // two OP_VARIFOLD sums
    double t[3];
    for (int k = 0; k < 3; k++) t[k] = 2.0 * (ktS[3 * f + k] - ktT[3 * f + k]);
    const double nt = n[0] * t[0] + n[1] * t[1] + n[2] * t[2];  // S(N N^T) t = N (N . t)
    const double r = 1.0 / sqrt(n2);
    for (int k = 0; k < 3; k++) kt[k] = (t[k] - n[k] * nt / (2 * n2)) * r;
  }
  for (int c = 0; c < 3; c++) {
    double x[3];
    cross3(e[c], kt, x);
    for (int k = 0; k < 3; k++) out[3 * c + k] = x[k] + gk[k];
  }
}

// grad[v] = scale * sum of the corner contributions (D values each) incident to v, in
// incidence order (nsub meshes sharing one incidence list of ncorner corners per mesh)
template <int D>
__global__ __launch_bounds__(256) void k_vertex_gather(const int *offsets, const int *inc,
                                                       long nv, long ncorner, int nsub,
                                                       const double *contrib, double scale,
                                                       double *grad) {
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv * nsub) return;
  const long sub = v / nv, vl = v - sub * nv;
  contrib += sub * ncorner * D;
  double g[D];
  for (int k = 0; k < D; k++) g[k] = 0.0;
  const long b = max(offsets[vl], 0), e = min((long)offsets[vl + 1], ncorner);
  for (long i = b; i < e; i++) {
    const int c = inc[i];
    if (c < 0 || c >= ncorner) continue;
    for (int k = 0; k < D; k++) g[k] += contrib[D * (long)c + k];
  }
  for (int k = 0; k < D; k++) grad[D * v + k] = scale * g[k];
}

// out[s] = sum over the rows of segment s of dot(X_row, Y_row) (dv columns); segments of
// seg_len rows, or [off[s], off[s + 1]) when off is given.  One workgroup per segment.
__global__ __launch_bounds__(256) void k_seg_dot(const double *X, const double *Y, int dv,
                                                 long seg_len, const long *off, double *out) {
  __shared__ double sh[256];
  const int sub = blockIdx.x;
  const long lo = off ? off[sub] : sub * seg_len;
  const long hi = off ? off[sub + 1] : lo + seg_len;
  double s = 0.0;
  for (long f = lo + threadIdx.x; f < hi; f += 256)
    for (int k = 0; k < dv; k++) s += X[dv * f + k] * Y[dv * f + k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[sub] = sh[0];
}

// Oriented samples (centre c, area-weighted normal N): kernel weights W and coefficients
// A as for faces; Nt = N for currents, N / sqrt(|N| + 1e-20) for varifolds.
__global__ __launch_bounds__(256) void k_sample_geometry(int kind, const double *N, long n,
                                                         double *W, double *A, double *Nt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[3] = {N[3 * i], N[3 * i + 1], N[3 * i + 2]};
  if (kind == DFM_CURRENT) {
    for (int k = 0; k < 3; k++) W[3 * i + k] = A[3 * i + k] = Nt[3 * i + k] = v[k];
    return;
  }
  const double s = sqrt(sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 1e-20);
  for (int k = 0; k < 3; k++) W[3 * i + k] = A[3 * i + k] = Nt[3 * i + k] = v[k] / s;
}

// d match / d c_i = 2 (g1_i - g2_i);  d match / d N_i = 2 (k1_i - k2_i) for currents and,
// for varifolds, the chain rule of Nt = N / s, s = sqrt(|N| + 1e-20), applied to
// d match / d Nt_i = 4 (S(k1_i) - S(k2_i)) Nt_i, S the symmetric matrix of a weight row:
//   d/dN = g / s - N (N.g) / (2 s^3 |N|)
__global__ __launch_bounds__(256) void k_sample_grads(int kind, const double *N,
                                                      const double *Nt, const double *k1,
                                                      const double *k2, const double *g1,
                                                      const double *g2, long n, double *grad_c,
                                                      double *grad_n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (grad_c)
    for (int k = 0; k < 3; k++) grad_c[3 * i + k] = 2.0 * (g1[3 * i + k] - g2[3 * i + k]);
  if (!grad_n) return;
  if (kind == DFM_CURRENT) {
    for (int k = 0; k < 3; k++) grad_n[3 * i + k] = 2.0 * (k1[3 * i + k] - k2[3 * i + k]);
    return;
  }
  // S((K W_S)(c_i) - (K W_T)(c_i)) u_i: the Convolve parts of the two OP_VARIFOLD sums
  double g[3];
  for (int k = 0; k < 3; k++) g[k] = k1[3 * i + k] - k2[3 * i + k];
  const double *v = N + 3 * i;
  const double a = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (!(a > 0.0)) {
    for (int k = 0; k < 3; k++) grad_n[3 * i + k] = 0.0;
    return;
  }
  const double s = sqrt(a + 1e-20);
  for (int k = 0; k < 3; k++) g[k] *= 4.0;
  const double ng = v[0] * g[0] + v[1] * g[1] + v[2] * g[2];
  for (int k = 0; k < 3; k++) grad_n[3 * i + k] = g[k] / s - v[k] * ng / (2.0 * s * s * s * a);
}

// ---- polylines: segments (i0, i1) of points in D = 2 or 3 ----------------------------

// Geometry of each segment: centre (p0 + p1) / 2; current W = A = T = p1 - p0
// (OrientedPolyLine.cxx:249-272); varifold W = A = T = (p1 - p0) / sqrt(|p1 - p0| + 1e-20)
// (NonOrientedPolyLine.cxx:266-297; its upper triangle of T T^T enters through OP_VARIFOLD).
// Bad indices flag the segment (weights zero).
// nsub polylines share one segment list E (ne segments over nv points): segment f of the
// output belongs to polyline f / ne, whose points start at row (f / ne) * nv of V.
template <int D>
__global__ __launch_bounds__(256) void k_seg_geometry(int kind, const double *V, long nv,
                                                      const int *E, long ne, int nsub, double *C,
                                                      double *W, double *A, double *T,
                                                      double *bad) {
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= ne * nsub) return;
  const long sub = f / ne, fl = f - sub * ne;
  V += sub * nv * D;
  const int i0 = E[2 * fl], i1 = E[2 * fl + 1];
  if (i0 < 0 || i1 < 0 || i0 >= nv || i1 >= nv) {
    for (int k = 0; k < D; k++) C[D * f + k] = 0.0, T[D * f + k] = 0.0;
    for (int k = 0; k < D; k++) W[D * f + k] = 0.0, A[D * f + k] = 0.0;
    *bad = 1.0;
    return;
  }
  double t[D], t2 = 0.0;
  for (int k = 0; k < D; k++) {
    const double a = V[D * (long)i0 + k], b = V[D * (long)i1 + k];
    C[D * f + k] = (a + b) / 2.0;
    t[k] = b - a;
    t2 += t[k] * t[k];
  }
  if (kind == DFM_CURRENT) {
    for (int k = 0; k < D; k++) W[D * f + k] = A[D * f + k] = T[D * f + k] = t[k];
    return;
  }
  const double s = sqrt(sqrt(t2) + 1e-20);
  for (int k = 0; k < D; k++) W[D * f + k] = A[D * f + k] = T[D * f + k] = t[k] / s;
}

// contrib[f] (2 endpoints x D) of the match gradient, from kt = (K W_S)(c_S) - (K W_T)(c_S)
// and g = Grad1 differences at the segment centres:
//   current   endpoint 0: -2 kt + g, endpoint 1: 2 kt + g (OrientedPolyLine.cxx:216-234)
//   varifold  u = 4 S(kt) T, u = (u - T (T.u) / (2 |T|^2)) / |T|; endpoints -u + g, u + g,
//             nothing where |T|^2 <= 1e-20 (NonOrientedPolyLine.cxx:215-251)
template <int D>
__global__ __launch_bounds__(256) void k_seg_contrib(int kind, const int *E, long ne, int nsub,
                                                     long nv, const double *T, const double *ktS,
                                                     const double *gS, const double *ktT,
                                                     const double *gT, double *contrib) {
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= ne * nsub) return;
  const long fl = f - (f / ne) * ne;
  double *out = contrib + 2 * D * f;
  for (int k = 0; k < 2 * D; k++) out[k] = 0.0;
  const int i0 = E[2 * fl], i1 = E[2 * fl + 1];
  if (i0 < 0 || i1 < 0 || i0 >= nv || i1 >= nv) return;
  double u[D], g[D];
  for (int k = 0; k < D; k++) g[k] = gS[D * f + k] - gT[D * f + k];
  if (kind == DFM_CURRENT) {
    for (int k = 0; k < D; k++) u[k] = 2.0 * (ktS[D * f + k] - ktT[D * f + k]);
  } else {
    const double *n = T + D * f;
    double n2 = 0.0;
    for (int k = 0; k < D; k++) n2 += n[k] * n[k];
    if (!(n2 > 1e-20)) return;
    // S(kt) T = V_S - V_T, the Convolve parts of the two OP_VARIFOLD sums
    for (int k = 0; k < D; k++) u[k] = ktS[D * f + k] - ktT[D * f + k];
    double nu = 0.0;
    for (int k = 0; k < D; k++) nu += n[k] * (u[k] *= 4.0);
    const double r = 1.0 / sqrt(n2);
    for (int k = 0; k < D; k++) u[k] = (u[k] - n[k] * nu / (2 * n2)) * r;
  }
  for (int k = 0; k < D; k++) {
    out[k] = -u[k] + g[k];
    out[D + k] = u[k] + g[k];
  }
}

// ---- volume meshes: cells of D + 1 points (triangles in 2D, tetrahedra in 3D) ------------

// Centre (p0 + ... + pD) / (D + 1) and volume |det(p1 - p0, ..., pD - p0)| / D! of each
// cell (OrientedVolumeMesh.cxx:305-334); bad indices flag the cell (weight zero).
template <int D>
__global__ __launch_bounds__(256) void k_vol_geometry(const double *V, long nv, const int *Cl,
                                                      long nc, double *C, double *W,
                                                      double *bad) {
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nc) return;
  int id[D + 1];
  bool ok = true;
  for (int c = 0; c <= D; c++) {
    id[c] = Cl[(D + 1) * f + c];
    ok = ok && id[c] >= 0 && id[c] < nv;
  }
  if (!ok) {
    for (int k = 0; k < D; k++) C[D * f + k] = 0.0;
    W[f] = 0.0;
    *bad = 1.0;
    return;
  }
  double p[D + 1][D];
  for (int c = 0; c <= D; c++)
    for (int k = 0; k < D; k++) p[c][k] = V[D * (long)id[c] + k];
  for (int k = 0; k < D; k++) {
    double s = p[0][k];
    for (int c = 1; c <= D; c++) s += p[c][k];
    C[D * f + k] = s / (D + 1.0);
  }
  double e[D][D];
  for (int r = 0; r < D; r++)
    for (int k = 0; k < D; k++) e[r][k] = p[r + 1][k] - p[0][k];
  double det;
  if constexpr (D == 3)
    det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
          e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
          e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
  else
    det = e[0][0] * e[1][1] - e[0][1] * e[1][0];
  W[f] = fabs(det) / (D == 3 ? 6.0 : 2.0);
}

// contrib[f] ((D + 1) corners x D) of the match gradient as OrientedVolumeMesh writes it
// (OrientedVolumeMesh.cxx:196-283): the determinant's cofactor vector of each corner times
// Ktau = (K V_S)(c_f) - (K V_T)(c_f), scaled by the reference's float 1/3 in 3D, plus the
// centre term (Grad1_S - Grad1_T)(c_f) / ((D + 1) / 2).  The cofactors carry no sign of the
// determinant: the derivative of the match for positively oriented cells.
template <int D>
__global__ __launch_bounds__(256) void k_vol_contrib(const double *V, long nv, const int *Cl,
                                                     long nc, const double *ktS, const double *gS,
                                                     const double *ktT, const double *gT,
                                                     double *contrib) {
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nc) return;
  double *out = contrib + (D + 1) * D * f;
  for (int k = 0; k < (D + 1) * D; k++) out[k] = 0.0;
  int id[D + 1];
  for (int c = 0; c <= D; c++) {
    id[c] = Cl[(D + 1) * f + c];
    if (id[c] < 0 || id[c] >= nv) return;
  }
  double p[D + 1][D];
  for (int c = 0; c <= D; c++)
    for (int k = 0; k < D; k++) p[c][k] = V[D * (long)id[c] + k];
  double v[D + 1][D];
  if constexpr (D == 3) {
    double a[3], b[3];
    auto sub = [&](int i, int j, double *o) {
      for (int k = 0; k < 3; k++) o[k] = p[i][k] - p[j][k];
    };
    sub(3, 1, a); sub(2, 1, b); cross3(a, b, v[0]);
    sub(2, 0, a); sub(3, 0, b); cross3(a, b, v[1]);
    sub(3, 0, a); sub(1, 0, b); cross3(a, b, v[2]);
    sub(1, 0, a); sub(2, 0, b); cross3(a, b, v[3]);
  } else {
    v[0][0] = p[1][1] - p[2][1]; v[0][1] = p[2][0] - p[1][0];
    v[1][0] = p[2][1] - p[0][1]; v[1][1] = p[0][0] - p[2][0];
    v[2][0] = p[0][1] - p[1][1]; v[2][1] = p[1][0] - p[0][0];
  }
  const double factor = D == 3 ? (double)(1.0f / 3.0f) : 1.0;
  const double ktau = ktS[f] - ktT[f];
  for (int c = 0; c <= D; c++)
    for (int k = 0; k < D; k++)
      out[D * c + k] = factor * v[c][k] * ktau + (gS[D * f + k] - gT[D * f + k]) / ((D + 1.0) / 2.0);
}

size_t a256(size_t v) { return (v + 255) & ~(size_t)255; }

struct SampleLayout {
  double *wS, *aS, *tS, *wT, *aT, *tT, *k1, *k2, *kT, *g1, *g2, *p1, *p2, *p3;
  void *pw;
  size_t pw_bytes, total;
  SampleLayout(char *base, int kind, int prec, long ns, long nt, bool gradient) {
    const int dv = 3;
    size_t off = 0;
    auto take = [&](size_t count) {
      double *p = base ? (double *)(base + off) : nullptr;
      off += a256(count * sizeof(double));
      return p;
    };
    wS = take(dv * ns); aS = take(dv * ns); tS = take(3 * ns);
    wT = take(dv * nt); aT = take(dv * nt); tT = take(3 * nt);
    k1 = take(dv * ns); k2 = take(dv * ns); kT = take(dv * nt);
    g1 = gradient ? take(3 * ns) : nullptr;
    g2 = gradient ? take(3 * ns) : nullptr;
    p1 = take(kRedBlocks); p2 = take(kRedBlocks); p3 = take(kRedBlocks);
    const int op = gradient ? data_grad_op(kind) : data_conv_op(kind);
    pw_bytes = std::max({pairsum_workspace(op, prec, 3, dv, ns, ns),
                         pairsum_workspace(op, prec, 3, dv, ns, nt),
                         pairsum_workspace(data_conv_op(kind), prec, 3, dv, nt, nt)});
    pw = base ? (void *)(base + off) : nullptr;
    off += a256(pw_bytes);
    total = off;
  }
};

struct Layout {
  double *cS, *wS, *aS, *nS, *cT, *wT, *aT, *nT;
  double *k1, *k2, *k3, *g1, *g2;  // conv / grad outputs
  double *p1, *p2, *p3;            // reduction partials
  double *contrib;
  double *bad;
  void *pw;
  size_t pw_bytes, total;
  Layout(char *base, int kind, int prec, long nfs, long nft, bool gradient) {
    const int dv = 3;
    size_t off = 0;
    auto take = [&](size_t count) {
      double *p = base ? (double *)(base + off) : nullptr;
      off += a256(count * sizeof(double));
      return p;
    };
    cS = take(3 * nfs); wS = take(dv * nfs); aS = take(dv * nfs); nS = take(3 * nfs);
    cT = take(3 * nft); wT = take(dv * nft); aT = take(dv * nft); nT = take(3 * nft);
    const long nmax = std::max(nfs, nft);
    k1 = take(dv * nmax); k2 = take(dv * nmax); k3 = take(dv * nmax);
    g1 = take(3 * nfs); g2 = take(3 * nfs);
    p1 = take(kRedBlocks); p2 = take(kRedBlocks); p3 = take(kRedBlocks);
    contrib = gradient ? take(9 * nfs) : nullptr;
    bad = take(1);
    pw_bytes = 0;
    if (gradient) {
      const int op = data_grad_op(kind);
      pw_bytes = std::max(pw_bytes, pairsum_workspace(op, prec, 3, dv, nfs, nfs));
      pw_bytes = std::max(pw_bytes, pairsum_workspace(op, prec, 3, dv, nfs, nft));
    } else {
      const int op = data_conv_op(kind);
      pw_bytes = std::max(pw_bytes, pairsum_workspace(op, prec, 3, dv, nfs, nfs));
      pw_bytes = std::max(pw_bytes, pairsum_workspace(op, prec, 3, dv, nft, nft));
      pw_bytes = std::max(pw_bytes, pairsum_workspace(op, prec, 3, dv, nft, nfs));
    }
    pw = base ? (void *)(base + off) : nullptr;
    off += a256(pw_bytes);
    total = off;
  }
};

struct CurveLayout {
  double *cS, *wS, *aS, *tS, *cT, *wT, *aT, *tT;
  double *k1, *k2, *k3, *g1, *g2;
  double *p1, *p2, *p3;
  double *contrib;
  double *bad;
  void *pw;
  size_t pw_bytes, total;
  CurveLayout(char *base, int kind, int D, int prec, long nes, long net, bool gradient) {
    const int dv = D;
    size_t off = 0;
    auto take = [&](size_t count) {
      double *p = base ? (double *)(base + off) : nullptr;
      off += a256(count * sizeof(double));
      return p;
    };
    cS = take(D * nes); wS = take(dv * nes); aS = take(dv * nes); tS = take(D * nes);
    cT = take(D * net); wT = take(dv * net); aT = take(dv * net); tT = take(D * net);
    const long nmax = std::max(nes, net);
    k1 = take(dv * nmax); k2 = take(dv * nmax); k3 = take(dv * nmax);
    g1 = take(D * nes); g2 = take(D * nes);
    p1 = take(kRedBlocks); p2 = take(kRedBlocks); p3 = take(kRedBlocks);
    contrib = gradient ? take(2 * D * nes) : nullptr;
    bad = take(1);
    if (gradient)
      pw_bytes = std::max(pairsum_workspace(data_grad_op(kind), prec, D, dv, nes, nes),
                          pairsum_workspace(data_grad_op(kind), prec, D, dv, nes, net));
    else
      pw_bytes = std::max({pairsum_workspace(data_conv_op(kind), prec, D, dv, nes, nes),
                           pairsum_workspace(data_conv_op(kind), prec, D, dv, net, net),
                           pairsum_workspace(data_conv_op(kind), prec, D, dv, net, nes)});
    pw = base ? (void *)(base + off) : nullptr;
    off += a256(pw_bytes);
    total = off;
  }
};

struct VolumeLayout {
  double *cS, *wS, *cT, *wT, *k1, *k2, *k3, *g1, *g2, *p1, *p2, *p3, *contrib, *bad;
  void *pw;
  size_t pw_bytes, total;
  VolumeLayout(char *base, int D, int prec, long ncs, long nct, bool gradient) {
    size_t off = 0;
    auto take = [&](size_t count) {
      double *p = base ? (double *)(base + off) : nullptr;
      off += a256(count * sizeof(double));
      return p;
    };
    cS = take(D * ncs); wS = take(ncs); cT = take(D * nct); wT = take(nct);
    const long nmax = std::max(ncs, nct);
    k1 = take(nmax); k2 = take(nmax); k3 = take(nmax);
    g1 = take(D * ncs); g2 = take(D * ncs);
    p1 = take(kRedBlocks); p2 = take(kRedBlocks); p3 = take(kRedBlocks);
    contrib = gradient ? take((D + 1) * D * ncs) : nullptr;
    bad = take(1);
    if (gradient)
      pw_bytes = std::max(pairsum_workspace(OP_CONVGRAD1, prec, D, 1, ncs, ncs),
                          pairsum_workspace(OP_CONVGRAD1, prec, D, 1, ncs, nct));
    else
      pw_bytes = std::max({pairsum_workspace(OP_CONV, prec, D, 1, ncs, ncs),
                           pairsum_workspace(OP_CONV, prec, D, 1, nct, nct),
                           pairsum_workspace(OP_CONV, prec, D, 1, nct, ncs)});
    pw = base ? (void *)(base + off) : nullptr;
    off += a256(pw_bytes);
    total = off;
  }
};

int launched(const char *what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

unsigned blocks_for(long n) { return (unsigned)std::max(1L, (n + 255) / 256); }

KParams blank() {
  KParams p;
  std::memset(&p, 0, sizeof(p));
  return p;
}

int check_args(int kind, int prec, double sigma, long nvs, long nfs, long nvt, long nft) {
  if (kind != DFM_CURRENT && kind != DFM_VARIFOLD)
    return set_error(DFM_ERR_ARG, "unknown data term %d", kind);
  if (!valid_precision(prec))
    return set_error(DFM_ERR_ARG, "unknown precision %d", prec);
  if (!(sigma == sigma)) return set_error(DFM_ERR_ARG, "kernel width is NaN");
  if (nvs < 0 || nfs < 0 || nvt < 0 || nft < 0 || nfs > 0x7fffffffL || nft > 0x7fffffffL ||
      nvs > 0x7fffffffL || nvt > 0x7fffffffL)
    return set_error(DFM_ERR_ARG, "counts must be in [0, 2^31)");
  return DFM_OK;
}

// geometry of both meshes and the reset of the bad-face flag
int geometry(hipStream_t st, int kind, const Layout &L, const double *vs, long nvs, const int *fs,
             long nfs, const double *vt, long nvt, const int *ft, long nft) {
  int rc = fill(st, L.bad, 1, 0.0);  // a fill kernel, so that callers may capture the call
  if (rc) return rc;
  if (nfs > 0)
    hipLaunchKernelGGL(k_face_geometry, dim3(blocks_for(nfs)), dim3(256), 0, st, kind, vs, nvs, fs,
                       nfs, 1, L.cS, L.wS, L.aS, L.nS, L.bad);
  if (nft > 0)
    hipLaunchKernelGGL(k_face_geometry, dim3(blocks_for(nft)), dim3(256), 0, st, kind, vt, nvt, ft,
                       nft, 1, L.cT, L.wT, L.aT, L.nT, L.bad);
  return launched("k_face_geometry");
}

int report_bad(hipStream_t st, const double *flag, const char *cell) {
  double bad = 0.0;
  hipError_t e = hipMemcpyAsync(&bad, flag, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "flag read: %s", hipGetErrorString(e));
  return bad != 0.0 ? set_error(DFM_ERR_ARG, "a %s references a vertex index outside [0, nv)", cell)
                    : DFM_OK;
}

int report_bad(hipStream_t st, const Layout &L) { return report_bad(st, L.bad, "face"); }

int check_curve_args(int kind, int dim, int prec, double sigma, long nvs, long nes, long nvt,
                     long net) {
  if (dim != 2 && dim != 3) return set_error(DFM_ERR_DIM, "polylines need dimension 2 or 3, got %d", dim);
  return check_args(kind, prec, sigma, nvs, nes, nvt, net);
}

template <int D>
int curve_geometry(hipStream_t st, int kind, const CurveLayout &L, const double *vs, long nvs,
                   const int *es, long nes, const double *vt, long nvt, const int *et, long net) {
  int rc = fill(st, L.bad, 1, 0.0);
  if (rc) return rc;
  if (nes > 0)
    hipLaunchKernelGGL(k_seg_geometry<D>, dim3(blocks_for(nes)), dim3(256), 0, st, kind, vs, nvs,
                       es, nes, 1, L.cS, L.wS, L.aS, L.tS, L.bad);
  if (net > 0)
    hipLaunchKernelGGL(k_seg_geometry<D>, dim3(blocks_for(net)), dim3(256), 0, st, kind, vt, nvt,
                       et, net, 1, L.cT, L.wT, L.aT, L.tT, L.bad);
  return launched("k_seg_geometry");
}

template <int D>
int curve_match(hipStream_t st, int kind, double sigma, int prec, const CurveLayout &L,
                const double *vs, long nvs, const int *es, long nes, const double *vt, long nvt,
                const int *et, long net, double *match) {
  const int dv = D, op = data_conv_op(kind);
  int rc = curve_geometry<D>(st, kind, L, vs, nvs, es, nes, vt, nvt, et, net);
  if (rc) return rc;
  KParams p = blank();
  // |S|^2 = sum_f A_S,f . (K W_S)(c_S,f) (OrientedPolyLine.cxx:279-296,
  // NonOrientedPolyLine.cxx:304-324), |T|^2 likewise, <S, T> = sum_f A_T,f . (K W_S)(c_T,f)
  // (the varifold's OP_VARIFOLD takes the targets' own u as A)
  p.x = L.cS; p.ta = L.aS; p.nx = nes; p.y = L.cS; p.w = L.wS; p.ny = nes; p.out0 = L.k1;
  if ((rc = pairsum(st, op, prec, D, dv, sigma, p, nes * dv, 0, L.pw, L.pw_bytes))) return rc;
  p.x = L.cT; p.ta = L.aT; p.nx = net; p.y = L.cT; p.w = L.wT; p.ny = net; p.out0 = L.k2;
  if ((rc = pairsum(st, op, prec, D, dv, sigma, p, net * dv, 0, L.pw, L.pw_bytes))) return rc;
  p.x = L.cT; p.ta = L.aT; p.nx = net; p.y = L.cS; p.w = L.wS; p.ny = nes; p.out0 = L.k3;
  if ((rc = pairsum(st, op, prec, D, dv, sigma, p, net * dv, 0, L.pw, L.pw_bytes))) return rc;
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k1, L.aS, nes, dv, L.p1);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k2, L.aT, net, dv, L.p2);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k3, L.aT, net, dv, L.p3);
  hipLaunchKernelGGL(k_combine, dim3(1), dim3(256), 0, st, L.p2, 1.0, L.p1, 1.0, L.p3, -2.0, match);
  if ((rc = launched("match reduction"))) return rc;
  return report_bad(st, L.bad, "segment");
}

template <int D>
int curve_match_gradient(hipStream_t st, int kind, double sigma, int prec, const CurveLayout &L,
                         const double *vs, long nvs, const int *es, long nes, const int *offsets,
                         const int *incidences, const double *vt, long nvt, const int *et,
                         long net, double *grad) {
  const int dv = D, op = data_grad_op(kind);
  int rc = curve_geometry<D>(st, kind, L, vs, nvs, es, nes, vt, nvt, et, net);
  if (rc) return rc;
  KParams p = blank();
  // (K W_S)(c_S) with ConvolveGradient(c_S, A_S), then the same from the target; for the
  // varifold (OP_VARIFOLD) the Grad1 part is the reference's contraction of the per-weight
  // gradients with T T^T (NonOrientedPolyLine.cxx:239-249)
  p.x = L.cS; p.ta = L.aS; p.nx = nes; p.y = L.cS; p.w = L.wS; p.ny = nes;
  p.out0 = L.k1; p.out1 = L.g1;
  if ((rc = pairsum(st, op, prec, D, dv, sigma, p, nes * dv, nes * D, L.pw, L.pw_bytes)))
    return rc;
  p.y = L.cT; p.w = L.wT; p.ny = net; p.out0 = L.k2; p.out1 = L.g2;
  if ((rc = pairsum(st, op, prec, D, dv, sigma, p, nes * dv, nes * D, L.pw, L.pw_bytes)))
    return rc;
  if (nes > 0)
    hipLaunchKernelGGL(k_seg_contrib<D>, dim3(blocks_for(nes)), dim3(256), 0, st, kind, es, nes,
                       1, nvs, L.tS, L.k1, L.g1, L.k2, L.g2, L.contrib);
  if (nvs > 0)
    hipLaunchKernelGGL(k_vertex_gather<D>, dim3(blocks_for(nvs)), dim3(256), 0, st, offsets,
                       incidences, nvs, 2 * nes, 1, L.contrib, 1.0, grad);
  if ((rc = launched("gradient assembly"))) return rc;
  return report_bad(st, L.bad, "segment");
}

int check_volume_args(int dim, int prec, double sigma, long nvs, long ncs, long nvt, long nct) {
  if (dim != 2 && dim != 3)
    return set_error(DFM_ERR_DIM, "volume meshes need dimension 2 or 3, got %d", dim);
  return check_args(DFM_CURRENT, prec, sigma, nvs, ncs, nvt, nct);
}

template <int D>
int volume_geometry(hipStream_t st, const VolumeLayout &L, const double *vs, long nvs,
                    const int *cs, long ncs, const double *vt, long nvt, const int *ct,
                    long nct) {
  int rc = fill(st, L.bad, 1, 0.0);
  if (rc) return rc;
  if (ncs > 0)
    hipLaunchKernelGGL(k_vol_geometry<D>, dim3(blocks_for(ncs)), dim3(256), 0, st, vs, nvs, cs,
                       ncs, L.cS, L.wS, L.bad);
  if (nct > 0)
    hipLaunchKernelGGL(k_vol_geometry<D>, dim3(blocks_for(nct)), dim3(256), 0, st, vt, nvt, ct,
                       nct, L.cT, L.wT, L.bad);
  return launched("k_vol_geometry");
}

template <int D>
int volume_match(hipStream_t st, double sigma, int prec, const VolumeLayout &L, const double *vs,
                 long nvs, const int *cs, long ncs, const double *vt, long nvt, const int *ct,
                 long nct, double *match) {
  int rc = volume_geometry<D>(st, L, vs, nvs, cs, ncs, vt, nvt, ct, nct);
  if (rc) return rc;
  // |S|^2 = sum_f V_f (K V_S)(c_f) (OrientedVolumeMesh.cxx:341-358), |T|^2 likewise,
  // <S, T> = sum_f V_T,f (K V_S)(c_T,f) (:142-149)
  KParams p = blank();
  p.x = L.cS; p.nx = ncs; p.y = L.cS; p.w = L.wS; p.ny = ncs; p.out0 = L.k1;
  if ((rc = pairsum(st, OP_CONV, prec, D, 1, sigma, p, ncs, 0, L.pw, L.pw_bytes))) return rc;
  p.x = L.cT; p.nx = nct; p.y = L.cT; p.w = L.wT; p.ny = nct; p.out0 = L.k2;
  if ((rc = pairsum(st, OP_CONV, prec, D, 1, sigma, p, nct, 0, L.pw, L.pw_bytes))) return rc;
  p.x = L.cT; p.nx = nct; p.y = L.cS; p.w = L.wS; p.ny = ncs; p.out0 = L.k3;
  if ((rc = pairsum(st, OP_CONV, prec, D, 1, sigma, p, nct, 0, L.pw, L.pw_bytes))) return rc;
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k1, L.wS, ncs, 1, L.p1);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k2, L.wT, nct, 1, L.p2);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k3, L.wT, nct, 1, L.p3);
  hipLaunchKernelGGL(k_combine, dim3(1), dim3(256), 0, st, L.p2, 1.0, L.p1, 1.0, L.p3, -2.0, match);
  if ((rc = launched("match reduction"))) return rc;
  return report_bad(st, L.bad, "cell");
}

template <int D>
int volume_match_gradient(hipStream_t st, double sigma, int prec, const VolumeLayout &L,
                          const double *vs, long nvs, const int *cs, long ncs,
                          const int *offsets, const int *incidences, const double *vt, long nvt,
                          const int *ct, long nct, double *grad) {
  int rc = volume_geometry<D>(st, L, vs, nvs, cs, ncs, vt, nvt, ct, nct);
  if (rc) return rc;
  // (K V)(c_S) and ConvolveGradient(c_S, V_S) from the template and from the target
  // (OrientedVolumeMesh.cxx:186-194)
  KParams p = blank();
  p.x = L.cS; p.ta = L.wS; p.nx = ncs; p.y = L.cS; p.w = L.wS; p.ny = ncs;
  p.out0 = L.k1; p.out1 = L.g1;
  if ((rc = pairsum(st, OP_CONVGRAD1, prec, D, 1, sigma, p, ncs, ncs * D, L.pw, L.pw_bytes)))
    return rc;
  p.y = L.cT; p.w = L.wT; p.ny = nct; p.out0 = L.k2; p.out1 = L.g2;
  if ((rc = pairsum(st, OP_CONVGRAD1, prec, D, 1, sigma, p, ncs, ncs * D, L.pw, L.pw_bytes)))
    return rc;
  if (ncs > 0)
    hipLaunchKernelGGL(k_vol_contrib<D>, dim3(blocks_for(ncs)), dim3(256), 0, st, vs, nvs, cs,
                       ncs, L.k1, L.g1, L.k2, L.g2, L.contrib);
  if (nvs > 0)
    hipLaunchKernelGGL(k_vertex_gather<D>, dim3(blocks_for(nvs)), dim3(256), 0, st, offsets,
                       incidences, nvs, (D + 1) * ncs, 1, L.contrib, 1.0, grad);
  if ((rc = launched("gradient assembly"))) return rc;
  return report_bad(st, L.bad, "cell");
}

// ---- point clouds and landmarks -----------------------------------------------------------

// out = 2 (g1 + g2): PointCloud's (grad_SdotS + grad_SdotT) * 2 (PointCloud.cxx:189)
__global__ __launch_bounds__(256) void k_twice_sum(const double *g1, const double *g2, long n,
                                                   double *out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    out[i] = (g1[i] + g2[i]) * 2.0;
}

// out = 2 (X - Y): Landmark::ComputeMatchGradient (Landmark.cxx:318-319)
__global__ __launch_bounds__(256) void k_twice_diff(const double *X, const double *Y, long n,
                                                    double *out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    out[i] = (X[i] - Y[i]) * 2.0;
}

// partial[b] = sum over the block's entries of (X - Y)^2
__global__ __launch_bounds__(256) void k_sqdiff_partial(const double *X, const double *Y, long n,
                                                        double *partial) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const double d = X[i] - Y[i];
    s += d * d;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out = sum of kRedBlocks partials, in a fixed order
__global__ __launch_bounds__(256) void k_total(const double *p, double *out) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  sh[t] = p[t];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) sh[t] += sh[t + st];
    __syncthreads();
  }
  if (t == 0) *out = sh[0];
}

// Point clouds are Diracs at the points weighted by one scalar each: the kernel sums take
// the caller's points and weights as they are, with no geometry pass.
struct PointLayout {
  double *k1 = nullptr, *k2 = nullptr, *k3 = nullptr, *g1 = nullptr, *g2 = nullptr;
  double *p1 = nullptr, *p2 = nullptr, *p3 = nullptr;
  void *pw;
  size_t pw_bytes, total;
  PointLayout(char *base, int D, int prec, long ns, long nt, bool gradient) {
    size_t off = 0;
    auto take = [&](size_t count) {
      double *p = base ? (double *)(base + off) : nullptr;
      off += a256(count * sizeof(double));
      return p;
    };
    if (gradient) {
      g1 = take(D * ns); g2 = take(D * ns);
      pw_bytes = std::max(pairsum_workspace(OP_GRAD1, prec, D, 1, ns, ns),
                          pairsum_workspace(OP_GRAD1, prec, D, 1, ns, nt));
    } else {
      k1 = take(ns); k2 = take(nt); k3 = take(nt);
      p1 = take(kRedBlocks); p2 = take(kRedBlocks); p3 = take(kRedBlocks);
      pw_bytes = std::max({pairsum_workspace(OP_CONV, prec, D, 1, ns, ns),
                           pairsum_workspace(OP_CONV, prec, D, 1, nt, nt),
                           pairsum_workspace(OP_CONV, prec, D, 1, nt, ns)});
    }
    pw = base ? (void *)(base + off) : nullptr;
    off += a256(pw_bytes);
    total = off;
  }
};

int check_point_args(int dim, int prec, double sigma, long ns, long nt) {
  if (dim != 2 && dim != 3)
    return set_error(DFM_ERR_DIM, "point clouds need dimension 2 or 3, got %d", dim);
  return check_args(DFM_CURRENT, prec, sigma, ns, ns, nt, nt);
}

int pointcloud_match(hipStream_t st, int D, double sigma, int prec, const PointLayout &L,
                     const double *xs, const double *ws, long ns, const double *xt,
                     const double *wt, long nt, double *match) {
  // |S|^2 = sum_i w_i (K w)(x_i) (UpdateSelfNorm, PointCloud.cxx:231-250), |T|^2 likewise,
  // <S, T> = sum_j wt_j (K w)(xt_j) (ComputeMatch, :137-143)
  KParams p = blank();
  int rc;
  p.x = xs; p.nx = ns; p.y = xs; p.w = ws; p.ny = ns; p.out0 = L.k1;
  if ((rc = pairsum(st, OP_CONV, prec, D, 1, sigma, p, ns, 0, L.pw, L.pw_bytes))) return rc;
  p.x = xt; p.nx = nt; p.y = xt; p.w = wt; p.ny = nt; p.out0 = L.k2;
  if ((rc = pairsum(st, OP_CONV, prec, D, 1, sigma, p, nt, 0, L.pw, L.pw_bytes))) return rc;
  p.x = xt; p.nx = nt; p.y = xs; p.w = ws; p.ny = ns; p.out0 = L.k3;
  if ((rc = pairsum(st, OP_CONV, prec, D, 1, sigma, p, nt, 0, L.pw, L.pw_bytes))) return rc;
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k1, ws, ns, 1, L.p1);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k2, wt, nt, 1, L.p2);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k3, wt, nt, 1, L.p3);
  hipLaunchKernelGGL(k_combine, dim3(1), dim3(256), 0, st, L.p2, 1.0, L.p1, 1.0, L.p3, -2.0, match);
  return launched("match reduction");
}

int pointcloud_match_gradient(hipStream_t st, int D, double sigma, int prec,
                              const PointLayout &L, const double *xs, const double *ws, long ns,
                              const double *xt, const double *wt, long nt, double *grad) {
  // ConvolveGradient(x, w) with the template's points and weights as sources, then with the
  // target's, summed and doubled as the reference writes it (PointCloud.cxx:179-189): the
  // cross term enters with +, so this is d/dx |S + T|^2, not d/dx |S - T|^2
  KParams p = blank();
  int rc;
  p.x = xs; p.ta = ws; p.nx = ns; p.y = xs; p.w = ws; p.ny = ns; p.out0 = L.g1;
  if ((rc = pairsum(st, OP_GRAD1, prec, D, 1, sigma, p, ns * D, 0, L.pw, L.pw_bytes))) return rc;
  p.y = xt; p.w = wt; p.ny = nt; p.out0 = L.g2;
  if ((rc = pairsum(st, OP_GRAD1, prec, D, 1, sigma, p, ns * D, 0, L.pw, L.pw_bytes))) return rc;
  if (ns > 0)
    hipLaunchKernelGGL(k_twice_sum, dim3(blocks_for(ns * D)), dim3(256), 0, st, L.g1, L.g2,
                       ns * D, grad);
  return launched("k_twice_sum");
}

}  // namespace

int surf_geometry(hipStream_t st, int kind, const double *V, long nv, const int *F, long nf,
                  int nsub, double *C, double *W, double *A, double *N, double *bad) {
  if (nf * nsub <= 0) return DFM_OK;
  hipLaunchKernelGGL(k_face_geometry, dim3(blocks_for(nf * nsub)), dim3(256), 0, st, kind, V, nv,
                     F, nf, nsub, C, W, A, N, bad);
  return launched("k_face_geometry");
}

int surf_contrib(hipStream_t st, int kind, const double *V, long nv, const int *F, long nf,
                 int nsub, const double *N, const double *ktS, const double *gS,
                 const double *ktT, const double *gT, double *contrib) {
  if (nf * nsub <= 0) return DFM_OK;
  hipLaunchKernelGGL(k_face_contrib, dim3(blocks_for(nf * nsub)), dim3(256), 0, st, kind, V, nv,
                     F, nf, nsub, N, ktS, gS, ktT, gT, contrib);
  return launched("k_face_contrib");
}

int surf_gather(hipStream_t st, const int *offsets, const int *inc, long nv, long nf, int nsub,
                const double *contrib, double scale, double *grad) {
  if (nv * nsub <= 0) return DFM_OK;
  hipLaunchKernelGGL(k_vertex_gather<3>, dim3(blocks_for(nv * nsub)), dim3(256), 0, st, offsets, inc,
                     nv, 3 * nf, nsub, contrib, scale, grad);
  return launched("k_vertex_gather");
}

int cells_geometry(hipStream_t st, int kind, int D, int corners, const double *V, long nv,
                   const int *cells, long nc, int nsub, double *C, double *W, double *A,
                   double *N, double *bad) {
  if (corners == 3) {
    if (D != 3) return set_error(DFM_ERR_DIM, "triangle meshes need dimension 3");
    return surf_geometry(st, kind, V, nv, cells, nc, nsub, C, W, A, N, bad);
  }
  if (corners != 2 || (D != 2 && D != 3))
    return set_error(DFM_ERR_DIM, "cells of %d points in dimension %d", corners, D);
  if (nc * nsub <= 0) return DFM_OK;
  if (D == 2)
    hipLaunchKernelGGL(k_seg_geometry<2>, dim3(blocks_for(nc * nsub)), dim3(256), 0, st, kind, V,
                       nv, cells, nc, nsub, C, W, A, N, bad);
  else
    hipLaunchKernelGGL(k_seg_geometry<3>, dim3(blocks_for(nc * nsub)), dim3(256), 0, st, kind, V,
                       nv, cells, nc, nsub, C, W, A, N, bad);
  return launched("k_seg_geometry");
}

int cells_contrib(hipStream_t st, int kind, int D, int corners, const double *V, long nv,
                  const int *cells, long nc, int nsub, const double *N, const double *ktS,
                  const double *gS, const double *ktT, const double *gT, double *contrib) {
  if (corners == 3) return surf_contrib(st, kind, V, nv, cells, nc, nsub, N, ktS, gS, ktT, gT, contrib);
  if (nc * nsub <= 0) return DFM_OK;
  if (D == 2)
    hipLaunchKernelGGL(k_seg_contrib<2>, dim3(blocks_for(nc * nsub)), dim3(256), 0, st, kind,
                       cells, nc, nsub, nv, N, ktS, gS, ktT, gT, contrib);
  else
    hipLaunchKernelGGL(k_seg_contrib<3>, dim3(blocks_for(nc * nsub)), dim3(256), 0, st, kind,
                       cells, nc, nsub, nv, N, ktS, gS, ktT, gT, contrib);
  return launched("k_seg_contrib");
}

int cells_gather(hipStream_t st, int D, int corners, const int *offsets, const int *inc, long nv,
                 long nc, int nsub, const double *contrib, double scale, double *grad) {
  if (corners == 3) return surf_gather(st, offsets, inc, nv, nc, nsub, contrib, scale, grad);
  if (nv * nsub <= 0) return DFM_OK;
  if (D == 2)
    hipLaunchKernelGGL(k_vertex_gather<2>, dim3(blocks_for(nv * nsub)), dim3(256), 0, st, offsets,
                       inc, nv, 2 * nc, nsub, contrib, scale, grad);
  else
    hipLaunchKernelGGL(k_vertex_gather<3>, dim3(blocks_for(nv * nsub)), dim3(256), 0, st, offsets,
                       inc, nv, 2 * nc, nsub, contrib, scale, grad);
  return launched("k_vertex_gather");
}

int seg_dot(hipStream_t st, const double *X, const double *Y, int dv, int nsub, long seg_len,
            const long *off, double *out) {
  if (nsub <= 0) return DFM_OK;
  hipLaunchKernelGGL(k_seg_dot, dim3((unsigned)nsub), dim3(256), 0, st, X, Y, dv, seg_len, off,
                     out);
  return launched("k_seg_dot");
}

}  // namespace dfm

using namespace dfm;

extern "C" {

int dfm_cell_incidence(const int *cells, long ncells, int corners, long nv, int *offsets,
                       int *incidences) {
  clear_error();
  if (corners < 1 || corners > 4) return set_error(DFM_ERR_ARG, "cells of %d corners", corners);
  if (ncells < 0 || nv < 0 || nv > 0x7fffffffL || ncells > 0x7fffffffL / corners)
    return set_error(DFM_ERR_ARG, "bad counts");
  if ((ncells > 0 && (!cells || !incidences)) || !offsets)
    return set_error(DFM_ERR_ARG, "NULL argument");
  const long ncorner = (long)corners * ncells;
  std::vector<int> cnt((size_t)nv + 1, 0);
  for (long i = 0; i < ncorner; i++) {
    const int v = cells[i];
    if (v < 0 || v >= nv)
      return set_error(DFM_ERR_ARG, "cell %ld references vertex %d outside [0, %ld)", i / corners,
                       v, nv);
    cnt[v + 1]++;
  }
  for (long v = 0; v < nv; v++) cnt[v + 1] += cnt[v];
  for (long v = 0; v <= nv; v++) offsets[v] = cnt[v];
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (long i = 0; i < ncorner; i++) incidences[pos[cells[i]]++] = (int)i;  // corners * cell + corner
  return DFM_OK;
}

int dfm_mesh_incidence(const int *faces, long nf, long nv, int *offsets, int *incidences) {
  return dfm_cell_incidence(faces, nf, 3, nv, offsets, incidences);
}

size_t dfm_curve_workspace(int kind, int dim, long nes, long net, int precision, int gradient) {
  if (check_curve_args(kind, dim, precision, 1.0, 0, nes, 0, net)) return 0;
  return CurveLayout(nullptr, kind, dim, precision, nes, net, gradient != 0).total;
}

int dfm_curve_match(void *stream, int kind, int dim, double sigma, int prec, const double *vs,
                    long nvs, const int *es, long nes, const double *vt, long nvt, const int *et,
                    long net, double *match, void *work, size_t work_bytes) {
  clear_error();
  int rc = check_curve_args(kind, dim, prec, sigma, nvs, nes, nvt, net);
  if (rc) return rc;
  if (!match || (nes > 0 && (!vs || !es)) || (net > 0 && (!vt || !et)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  CurveLayout need(nullptr, kind, dim, prec, nes, net, false);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  CurveLayout L((char *)work, kind, dim, prec, nes, net, false);
  hipStream_t st = (hipStream_t)stream;
  return dim == 2 ? curve_match<2>(st, kind, sigma, prec, L, vs, nvs, es, nes, vt, nvt, et, net, match)
                  : curve_match<3>(st, kind, sigma, prec, L, vs, nvs, es, nes, vt, nvt, et, net, match);
}

int dfm_curve_match_gradient(void *stream, int kind, int dim, double sigma, int prec,
                             const double *vs, long nvs, const int *es, long nes,
                             const int *offsets, const int *incidences, const double *vt,
                             long nvt, const int *et, long net, double *grad, void *work,
                             size_t work_bytes) {
  clear_error();
  int rc = check_curve_args(kind, dim, prec, sigma, nvs, nes, nvt, net);
  if (rc) return rc;
  if (!grad || !offsets || (nes > 0 && (!vs || !es || !incidences)) || (net > 0 && (!vt || !et)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  CurveLayout need(nullptr, kind, dim, prec, nes, net, true);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  CurveLayout L((char *)work, kind, dim, prec, nes, net, true);
  hipStream_t st = (hipStream_t)stream;
  return dim == 2 ? curve_match_gradient<2>(st, kind, sigma, prec, L, vs, nvs, es, nes, offsets,
                                            incidences, vt, nvt, et, net, grad)
                  : curve_match_gradient<3>(st, kind, sigma, prec, L, vs, nvs, es, nes, offsets,
                                            incidences, vt, nvt, et, net, grad);
}

size_t dfm_surface_workspace(int kind, long nfs, long nft, int precision, int gradient) {
  if (check_args(kind, precision, 1.0, 0, nfs, 0, nft)) return 0;
  return Layout(nullptr, kind, precision, nfs, nft, gradient != 0).total;
}

int dfm_surface_match(void *stream, int kind, double sigma, int prec, const double *vs, long nvs,
                      const int *fs, long nfs, const double *vt, long nvt, const int *ft, long nft,
                      double *match, void *work, size_t work_bytes) {
  clear_error();
  int rc = check_args(kind, prec, sigma, nvs, nfs, nvt, nft);
  if (rc) return rc;
  if (!match || (nfs > 0 && (!vs || !fs)) || (nft > 0 && (!vt || !ft)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  Layout need(nullptr, kind, prec, nfs, nft, false);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  Layout L((char *)work, kind, prec, nfs, nft, false);
  hipStream_t st = (hipStream_t)stream;
  const int dv = 3, op = data_conv_op(kind);
  if ((rc = geometry(st, kind, L, vs, nvs, fs, nfs, vt, nvt, ft, nft))) return rc;
  KParams p = blank();
  // |S|^2: (K W_S)(c_S) . A_S (the varifold's OP_VARIFOLD takes the targets' own u as A)
  p.x = L.cS; p.ta = L.aS; p.nx = nfs; p.y = L.cS; p.w = L.wS; p.ny = nfs; p.out0 = L.k1;
  if ((rc = pairsum(st, op, prec, 3, dv, sigma, p, nfs * dv, 0, L.pw, L.pw_bytes))) return rc;
  // |T|^2
  p.x = L.cT; p.ta = L.aT; p.nx = nft; p.y = L.cT; p.w = L.wT; p.ny = nft; p.out0 = L.k2;
  if ((rc = pairsum(st, op, prec, 3, dv, sigma, p, nft * dv, 0, L.pw, L.pw_bytes))) return rc;
  // <S, T>: (K W_S)(c_T) . A_T
  p.x = L.cT; p.ta = L.aT; p.nx = nft; p.y = L.cS; p.w = L.wS; p.ny = nfs; p.out0 = L.k3;
  if ((rc = pairsum(st, op, prec, 3, dv, sigma, p, nft * dv, 0, L.pw, L.pw_bytes))) return rc;
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k1, L.aS, nfs, dv, L.p1);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k2, L.aT, nft, dv, L.p2);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k3, L.aT, nft, dv, L.p3);
  hipLaunchKernelGGL(k_combine, dim3(1), dim3(256), 0, st, L.p2, 1.0, L.p1, 1.0, L.p3, -2.0, match);
  if ((rc = launched("match reduction"))) return rc;
  return report_bad(st, L);
}

int dfm_surface_match_gradient(void *stream, int kind, double sigma, int prec, const double *vs,
                               long nvs, const int *fs, long nfs, const int *offsets,
                               const int *incidences, const double *vt, long nvt, const int *ft,
                               long nft, double *grad, void *work, size_t work_bytes) {
  clear_error();
  int rc = check_args(kind, prec, sigma, nvs, nfs, nvt, nft);
  if (rc) return rc;
  if (!grad || !offsets || (nfs > 0 && (!vs || !fs || !incidences)) || (nft > 0 && (!vt || !ft)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  Layout need(nullptr, kind, prec, nfs, nft, true);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  Layout L((char *)work, kind, prec, nfs, nft, true);
  hipStream_t st = (hipStream_t)stream;
  const int dv = 3, op = data_grad_op(kind);
  if ((rc = geometry(st, kind, L, vs, nvs, fs, nfs, vt, nvt, ft, nft))) return rc;
  KParams p = blank();
  // template on template: (K W_S)(c_S) and ConvolveGradient(c_S, A_S)
  p.x = L.cS; p.ta = L.aS; p.nx = nfs; p.y = L.cS; p.w = L.wS; p.ny = nfs;
  p.out0 = L.k1; p.out1 = L.g1;
  if ((rc = pairsum(st, op, prec, 3, dv, sigma, p, nfs * dv, nfs * 3, L.pw, L.pw_bytes)))
    return rc;
  // target on template
  p.y = L.cT; p.w = L.wT; p.ny = nft; p.out0 = L.k2; p.out1 = L.g2;
  if ((rc = pairsum(st, op, prec, 3, dv, sigma, p, nfs * dv, nfs * 3, L.pw, L.pw_bytes)))
    return rc;
  if (nfs > 0)
    hipLaunchKernelGGL(k_face_contrib, dim3(blocks_for(nfs)), dim3(256), 0, st, kind, vs, nvs, fs,
                       nfs, 1, L.nS, L.k1, L.g1, L.k2, L.g2, L.contrib);
  if (nvs > 0)
    hipLaunchKernelGGL(k_vertex_gather<3>, dim3(blocks_for(nvs)), dim3(256), 0, st, offsets,
                       incidences, nvs, 3 * nfs, 1, L.contrib, 1.0, grad);
  if ((rc = launched("gradient assembly"))) return rc;
  return report_bad(st, L);
}

size_t dfm_volume_workspace(int dim, long ncs, long nct, int precision, int gradient) {
  if (check_volume_args(dim, precision, 1.0, 0, ncs, 0, nct)) return 0;
  return VolumeLayout(nullptr, dim, precision, ncs, nct, gradient != 0).total;
}

int dfm_volume_match(void *stream, int dim, double sigma, int prec, const double *vs, long nvs,
                     const int *cs, long ncs, const double *vt, long nvt, const int *ct, long nct,
                     double *match, void *work, size_t work_bytes) {
  clear_error();
  int rc = check_volume_args(dim, prec, sigma, nvs, ncs, nvt, nct);
  if (rc) return rc;
  if (!match || (ncs > 0 && (!vs || !cs)) || (nct > 0 && (!vt || !ct)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  VolumeLayout need(nullptr, dim, prec, ncs, nct, false);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  VolumeLayout L((char *)work, dim, prec, ncs, nct, false);
  hipStream_t st = (hipStream_t)stream;
  return dim == 2 ? volume_match<2>(st, sigma, prec, L, vs, nvs, cs, ncs, vt, nvt, ct, nct, match)
                  : volume_match<3>(st, sigma, prec, L, vs, nvs, cs, ncs, vt, nvt, ct, nct, match);
}

int dfm_volume_match_gradient(void *stream, int dim, double sigma, int prec, const double *vs,
                              long nvs, const int *cs, long ncs, const int *offsets,
                              const int *incidences, const double *vt, long nvt, const int *ct,
                              long nct, double *grad, void *work, size_t work_bytes) {
  clear_error();
  int rc = check_volume_args(dim, prec, sigma, nvs, ncs, nvt, nct);
  if (rc) return rc;
  if (!grad || !offsets || (ncs > 0 && (!vs || !cs || !incidences)) || (nct > 0 && (!vt || !ct)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  VolumeLayout need(nullptr, dim, prec, ncs, nct, true);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  VolumeLayout L((char *)work, dim, prec, ncs, nct, true);
  hipStream_t st = (hipStream_t)stream;
  return dim == 2 ? volume_match_gradient<2>(st, sigma, prec, L, vs, nvs, cs, ncs, offsets,
                                             incidences, vt, nvt, ct, nct, grad)
                  : volume_match_gradient<3>(st, sigma, prec, L, vs, nvs, cs, ncs, offsets,
                                             incidences, vt, nvt, ct, nct, grad);
}

size_t dfm_pointcloud_workspace(int dim, long ns, long nt, int precision, int gradient) {
  if (check_point_args(dim, precision, 1.0, ns, nt)) return 0;
  return PointLayout(nullptr, dim, precision, ns, nt, gradient != 0).total;
}

int dfm_pointcloud_match(void *stream, int dim, double sigma, int prec, const double *xs,
                         const double *ws, long ns, const double *xt, const double *wt, long nt,
                         double *match, void *work, size_t work_bytes) {
  clear_error();
  int rc = check_point_args(dim, prec, sigma, ns, nt);
  if (rc) return rc;
  if (!match || (ns > 0 && (!xs || !ws)) || (nt > 0 && (!xt || !wt)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  PointLayout need(nullptr, dim, prec, ns, nt, false);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  PointLayout L((char *)work, dim, prec, ns, nt, false);
  return pointcloud_match((hipStream_t)stream, dim, sigma, prec, L, xs, ws, ns, xt, wt, nt,
                          match);
}

int dfm_pointcloud_match_gradient(void *stream, int dim, double sigma, int prec,
                                  const double *xs, const double *ws, long ns, const double *xt,
                                  const double *wt, long nt, double *grad, void *work,
                                  size_t work_bytes) {
  clear_error();
  int rc = check_point_args(dim, prec, sigma, ns, nt);
  if (rc) return rc;
  if ((ns > 0 && (!xs || !ws || !grad)) || (nt > 0 && (!xt || !wt)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  PointLayout need(nullptr, dim, prec, ns, nt, true);
  if (need.total > 0 && (!work || work_bytes < need.total))  // an empty template needs none
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  PointLayout L((char *)work, dim, prec, ns, nt, true);
  return pointcloud_match_gradient((hipStream_t)stream, dim, sigma, prec, L, xs, ws, ns, xt, wt,
                                   nt, grad);
}

size_t dfm_landmark_workspace(int dim, long n) {
  if ((dim != 2 && dim != 3) || n < 0 || n > 0x7fffffffL) return 0;
  return a256(kRedBlocks * sizeof(double));
}

int dfm_landmark_match(void *stream, int dim, const double *x, const double *y, long n,
                       double *match, void *work, size_t work_bytes) {
  clear_error();
  if (dim != 2 && dim != 3)
    return set_error(DFM_ERR_DIM, "landmarks need dimension 2 or 3, got %d", dim);
  if (n < 0 || n > 0x7fffffffL) return set_error(DFM_ERR_ARG, "counts must be in [0, 2^31)");
  if (!match || (n > 0 && (!x || !y))) return set_error(DFM_ERR_ARG, "NULL argument");
  if (!work || work_bytes < a256(kRedBlocks * sizeof(double)))
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes",
                     a256(kRedBlocks * sizeof(double)));
  hipStream_t st = (hipStream_t)stream;
  // |X - Y|_F^2 (Landmark.cxx:298-303), a deterministic two-stage FP64 sum
  hipLaunchKernelGGL(k_sqdiff_partial, dim3(kRedBlocks), dim3(256), 0, st, x, y, n * dim,
                     (double *)work);
  hipLaunchKernelGGL(k_total, dim3(1), dim3(256), 0, st, (const double *)work, match);
  return launched("landmark match");
}

int dfm_landmark_match_gradient(void *stream, int dim, const double *x, const double *y, long n,
                                double *grad) {
  clear_error();
  if (dim != 2 && dim != 3)
    return set_error(DFM_ERR_DIM, "landmarks need dimension 2 or 3, got %d", dim);
  if (n < 0 || n > 0x7fffffffL) return set_error(DFM_ERR_ARG, "counts must be in [0, 2^31)");
  if (n == 0) return DFM_OK;
  if (!x || !y || !grad) return set_error(DFM_ERR_ARG, "NULL argument");
  hipLaunchKernelGGL(k_twice_diff, dim3(blocks_for(n * dim)), dim3(256), 0, (hipStream_t)stream,
                     x, y, n * dim, grad);
  return launched("landmark gradient");
}

size_t dfm_samples_workspace(int kind, long ns, long nt, int precision, int gradient) {
  if (check_args(kind, precision, 1.0, 0, ns, 0, nt)) return 0;
  return SampleLayout(nullptr, kind, precision, ns, nt, gradient != 0).total;
}

int dfm_samples_match_gradient(void *stream, int kind, double sigma, int prec, const double *cs,
                               const double *Ns, long ns, const double *ct, const double *Nt,
                               long nt, double *match, double *grad_centers,
                               double *grad_normals, void *work, size_t work_bytes) {
  clear_error();
  int rc = check_args(kind, prec, sigma, 0, ns, 0, nt);
  if (rc) return rc;
  if (!match || (ns > 0 && (!cs || !Ns)) || (nt > 0 && (!ct || !Nt)))
    return set_error(DFM_ERR_ARG, "NULL argument");
  const bool gradient = grad_centers || grad_normals;
  SampleLayout need(nullptr, kind, prec, ns, nt, gradient);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  SampleLayout L((char *)work, kind, prec, ns, nt, gradient);
  hipStream_t st = (hipStream_t)stream;
  const int dv = 3;
  if (ns > 0)
    hipLaunchKernelGGL(k_sample_geometry, dim3(blocks_for(ns)), dim3(256), 0, st, kind, Ns, ns,
                       L.wS, L.aS, L.tS);
  if (nt > 0)
    hipLaunchKernelGGL(k_sample_geometry, dim3(blocks_for(nt)), dim3(256), 0, st, kind, Nt, nt,
                       L.wT, L.aT, L.tT);
  if ((rc = launched("k_sample_geometry"))) return rc;
  KParams p = blank();
  const int op = gradient ? data_grad_op(kind) : data_conv_op(kind);
  // (K W_S)(c_S) [+ Grad1(c_S, A_S)] and (K W_T)(c_S) [+ Grad1(c_S, A_S)]
  p.x = cs; p.ta = L.aS; p.nx = ns; p.y = cs; p.w = L.wS; p.ny = ns; p.out0 = L.k1; p.out1 = L.g1;
  if ((rc = pairsum(st, op, prec, 3, dv, sigma, p, ns * dv, gradient ? ns * 3 : 0, L.pw,
                    L.pw_bytes)))
    return rc;
  p.y = ct; p.w = L.wT; p.ny = nt; p.out0 = L.k2; p.out1 = L.g2;
  if ((rc = pairsum(st, op, prec, 3, dv, sigma, p, ns * dv, gradient ? ns * 3 : 0, L.pw,
                    L.pw_bytes)))
    return rc;
  // (K W_T)(c_T): the target's squared norm
  p = blank();
  p.x = ct; p.ta = L.aT; p.nx = nt; p.y = ct; p.w = L.wT; p.ny = nt; p.out0 = L.kT;
  if ((rc = pairsum(st, data_conv_op(kind), prec, 3, dv, sigma, p, nt * dv, 0, L.pw,
                    L.pw_bytes)))
    return rc;
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k1, L.aS, ns, dv, L.p1);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.kT, L.aT, nt, dv, L.p2);
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, st, L.k2, L.aS, ns, dv, L.p3);
  hipLaunchKernelGGL(k_combine, dim3(1), dim3(256), 0, st, L.p1, 1.0, L.p2, 1.0, L.p3, -2.0, match);
  if (gradient && ns > 0)
    hipLaunchKernelGGL(k_sample_grads, dim3(blocks_for(ns)), dim3(256), 0, st, kind, Ns, L.tS,
                       L.k1, L.k2, L.g1, L.g2, ns, grad_centers, grad_normals);
  return launched("samples match");
}

}  // extern "C"
