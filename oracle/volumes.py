"""ORACLE — test infrastructure only.

FP64 CPU restatement of Deformetrica's OrientedVolumeMesh data term (triangles in 2D,
tetrahedra in 3D) on top of the oracle ExactKernel:

  geometry  centres (p0 + ... + pD) / (D + 1) and volumes |det(p1 - p0, ..., pD - p0)| / D!
            with D! = 2 or 6 (OrientedVolumeMesh.cxx:305-334)
  norm      sum_f V_f (K V)(c_f) (:341-358)
  match     |S|^2 + |T|^2 - 2 <S, T> (:116-152)
  gradient  (:158-285) as the reference writes it: per cell, the cofactor vectors v_c of the
            determinant (cross products in 3D, rotated edges in 2D) times
            Ktau = (K V_S)(c_f) - (K V_T)(c_f), with the factor float(1/3) in 3D (the
            reference's 1.0f/3.0f) and 1 in 2D, plus the centre term
            (Grad1_S - Grad1_T)(c_f) / ((D + 1) / 2) at every corner.  The cofactors carry
            no sign of the determinant, so the formula is the derivative of the match for
            positively oriented cells (det > 0), which the reference's meshes are.

Pinned by tests/test_volumes.py: the match against a brute-force double sum over cell
pairs, the gradient against central differences on positively oriented meshes, and the
reference's own check (TestVolumeGradient.cxx:79-112: forward differences, h = 1e-5, within
1e-5) on its volume meshes.
"""
from __future__ import annotations

import numpy as np

from .oracle import ExactKernel

THIRD_F32 = float(np.float32(1.0) / np.float32(3.0))


def volume_geometry(V, C):
    V = np.asarray(V, dtype=np.float64)
    C = np.asarray(C)
    D = V.shape[1]
    P = V[C]                                     # cells x (D + 1) x D
    centre = P[:, 0].copy()
    for k in range(1, D + 1):
        centre += P[:, k]
    centre /= (D + 1.0)
    edges = P[:, 1:] - P[:, :1]                  # rows p_k - p0
    vol = np.abs(np.linalg.det(edges)) / (6.0 if D == 3 else 2.0)
    return centre, vol[:, None]


def _kernel(Y, W, width):
    k = ExactKernel(Y.shape[1], width)
    k.SetSources(Y)
    k.SetWeights(W)
    return k


def volume_norm2(V, C, width):
    c, w = volume_geometry(V, C)
    return float(np.sum(_kernel(c, w, width).Convolve(c) * w))


def volume_match(Vs, Cs, Vt, Ct, width):
    cs, ws = volume_geometry(Vs, Cs)
    ct, wt = volume_geometry(Vt, Ct)
    sdot = _kernel(cs, ws, width).Convolve(ct)
    return volume_norm2(Vt, Ct, width) + volume_norm2(Vs, Cs, width) - 2.0 * float(np.sum(sdot * wt))


def _cofactors(P):
    """d det / d p_c for one cell P ((D + 1) x D), c = 0..D."""
    if P.shape[1] == 3:
        p0, p1, p2, p3 = P
        return [np.cross(p3 - p1, p2 - p1), np.cross(p2 - p0, p3 - p0),
                np.cross(p3 - p0, p1 - p0), np.cross(p1 - p0, p2 - p0)]
    p0, p1, p2 = P
    return [np.array([p1[1] - p2[1], p2[0] - p1[0]]), np.array([p2[1] - p0[1], p0[0] - p2[0]]),
            np.array([p0[1] - p1[1], p1[0] - p0[0]])]


def volume_match_gradient(Vs, Cs, Vt, Ct, width):
    Vs = np.asarray(Vs, dtype=np.float64)
    D = Vs.shape[1]
    cs, ws = volume_geometry(Vs, Cs)
    ct, wt = volume_geometry(Vt, Ct)
    kS = _kernel(cs, ws, width)
    ktS, gS = kS.Convolve(cs)[:, 0], kS.ConvolveGradient(cs, ws)
    kT = _kernel(ct, wt, width)
    ktT, gT = kT.Convolve(cs)[:, 0], kT.ConvolveGradient(cs, ws)
    gk = (gS - gT) / ((D + 1.0) / 2.0)
    factor = THIRD_F32 if D == 3 else 1.0
    grad = np.zeros_like(Vs)
    for f, cell in enumerate(np.asarray(Cs)):
        ktau = ktS[f] - ktT[f]
        for c, v in zip(cell, _cofactors(Vs[cell])):
            grad[c] += factor * v * ktau
            grad[c] += gk[f]
    return grad


def brute_force_volume(Vs, Cs, Vt, Ct, width):
    """|S - T|^2 of the two volume measures as an explicit double sum over cells."""
    cs, ws = volume_geometry(Vs, Cs)
    ct, wt = volume_geometry(Vt, Ct)
    c = np.concatenate([cs, ct])
    w = np.concatenate([ws[:, 0], -wt[:, 0]])
    r2 = np.sum((c[:, None, :] - c[None, :, :]) ** 2, axis=-1)
    return float(np.sum(np.exp(-r2 / width ** 2) * np.outer(w, w)))
