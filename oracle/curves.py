"""ORACLE — test infrastructure only.

FP64 CPU restatement of Deformetrica's polyline data terms (2D or 3D segments) on top of
the oracle ExactKernel:

  current   OrientedPolyLine: centres (p0 + p1) / 2 and tangents p1 - p0
            (OrientedPolyLine.cxx:249-272), self norm (:279-296), ComputeMatch (:132-170),
            ComputeMatchGradient (:177-238)
  varifold  NonOrientedPolyLine: tangents scaled by |t|^-1/2 and the upper-triangular T T^T
            weights (NonOrientedPolyLine.cxx:266-297), self norm (:304-324), special_product
            (:358-370), ComputeMatch (:129-170), ComputeMatchGradient (:177-255)

The loops follow the reference segment by segment.  Pinned by tests/test_curves.py: each
match against a brute-force double sum over segment pairs, each gradient against central
differences of its match.
"""
from __future__ import annotations

import numpy as np

from .oracle import ExactKernel


def _kernel(Y, W, width):
    k = ExactKernel(Y.shape[1], width)
    k.SetSources(Y)
    k.SetWeights(W)
    return k


def _ends(V, E):
    E = np.asarray(E)
    return V[E[:, 0]], V[E[:, 1]]


def current_geometry(V, E):
    p0, p1 = _ends(V, E)
    return (p0 + p1) / 2.0, p1 - p0


def sym_upper(T):
    """Upper triangle (p <= q, row by row) of T T^T, per row of T."""
    d = T.shape[1]
    return np.stack([T[:, p] * T[:, q] for p in range(d) for q in range(p, d)], axis=1)


def varifold_geometry(V, E):
    p0, p1 = _ends(V, E)
    t = p1 - p0
    T = t / np.sqrt(np.linalg.norm(t, axis=1, keepdims=True) + 1e-20)
    return (p0 + p1) / 2.0, T, sym_upper(T)


def special_product(M, X):
    """The symmetric matrix whose upper triangle is M, times X, row by row."""
    d = X.shape[-1]
    S = np.empty(M.shape[:-1] + (d, d))
    idx = 0
    for p in range(d):
        for q in range(p, d):
            S[..., p, q] = M[..., idx]
            S[..., q, p] = M[..., idx]
            idx += 1
    return np.einsum("...pq,...q->...p", S, X)


def current_norm2(V, E, width):
    c, t = current_geometry(V, E)
    return float(np.sum(_kernel(c, t, width).Convolve(c) * t))


def current_match(Vs, Es, Vt, Et, width):
    cs, ts = current_geometry(Vs, Es)
    ct, tt = current_geometry(Vt, Et)
    sdot = _kernel(cs, ts, width).Convolve(ct)
    return current_norm2(Vt, Et, width) + current_norm2(Vs, Es, width) - 2.0 * float(np.sum(sdot * tt))


def current_match_gradient(Vs, Es, Vt, Et, width):
    cs, ts = current_geometry(Vs, Es)
    ct, tt = current_geometry(Vt, Et)
    kS = _kernel(cs, ts, width)
    ktS, gS = kS.Convolve(cs), kS.ConvolveGradient(cs, ts)
    kT = _kernel(ct, tt, width)
    ktT, gT = kT.Convolve(cs), kT.ConvolveGradient(cs, ts)
    gk = gS - gT
    grad = np.zeros_like(Vs)
    for f, (i0, i1) in enumerate(np.asarray(Es)):
        kt = 2.0 * (ktS[f] - ktT[f])
        grad[i0] += -kt + gk[f]
        grad[i1] += kt + gk[f]
    return grad


def varifold_norm2(V, E, width):
    c, T, M = varifold_geometry(V, E)
    kw = _kernel(c, M, width).Convolve(c)
    return float(np.sum(T * special_product(kw, T)))


def varifold_match(Vs, Es, Vt, Et, width):
    cs, Ts, Ms = varifold_geometry(Vs, Es)
    ct, Tt, Mt = varifold_geometry(Vt, Et)
    sdot = _kernel(cs, Ms, width).Convolve(ct)
    cross = float(np.sum(Tt * special_product(sdot, Tt)))
    return varifold_norm2(Vt, Et, width) + varifold_norm2(Vs, Es, width) - 2.0 * cross


def varifold_match_gradient(Vs, Es, Vt, Et, width):
    d = Vs.shape[1]
    cs, Ts, Ms = varifold_geometry(Vs, Es)
    ct, Tt, Mt = varifold_geometry(Vt, Et)
    kS = _kernel(cs, Ms, width)
    ktS, gS = kS.Convolve(cs), kS.ConvolveGradient(cs)   # (F, m), (F, m, d)
    kT = _kernel(ct, Mt, width)
    ktT, gT = kT.Convolve(cs), kT.ConvolveGradient(cs)
    grad = np.zeros_like(Vs)
    for f, (i0, i1) in enumerate(np.asarray(Es)):
        n = Ts[f]
        n2 = float(n @ n)
        if n2 > 1e-20:
            kt = 4.0 * special_product(ktS[f] - ktT[f], n)
            aux = special_product(Ms[f], kt)
            kt = (kt - aux / (2 * n2)) / np.sqrt(n2)
            grad[i0] -= kt
            grad[i1] += kt
            delta = gS[f] - gT[f]
            gk = np.array([special_product(delta[:, p], n) @ n for p in range(d)])
            grad[i0] += gk
            grad[i1] += gk
    return grad


def brute_force_current(Vs, Es, Vt, Et, width):
    """|S - T|^2 of the two segment currents as an explicit double sum."""
    cs, ts = current_geometry(Vs, Es)
    ct, tt = current_geometry(Vt, Et)
    c = np.concatenate([cs, ct])
    t = np.concatenate([ts, -tt])
    r2 = np.sum((c[:, None, :] - c[None, :, :]) ** 2, axis=-1)
    return float(np.sum(np.exp(-r2 / width ** 2) * (t @ t.T)))


def brute_force_varifold(Vs, Es, Vt, Et, width):
    """|S - T|^2 of the two segment varifolds, sum K (T_i . T_j)^2 with signs."""
    cs, Ts, _ = varifold_geometry(Vs, Es)
    ct, Tt, _ = varifold_geometry(Vt, Et)
    c = np.concatenate([cs, ct])
    T = np.concatenate([Ts, Tt])
    sgn = np.concatenate([np.ones(len(cs)), -np.ones(len(ct))])
    r2 = np.sum((c[:, None, :] - c[None, :, :]) ** 2, axis=-1)
    return float(np.sum(np.exp(-r2 / width ** 2) * (T @ T.T) ** 2 * np.outer(sgn, sgn)))
