"""ORACLE — test infrastructure only.

FP64 CPU restatement of Deformetrica's triangle-surface data terms on top of the oracle
ExactKernel:

  current   OrientedSurfaceMesh: centers/normals (OrientedSurfaceMesh.cxx:298-326),
            self norm (:330-351), ComputeMatch (:125-163), ComputeMatchGradient (:166-237)
  varifold  NonOrientedSurfaceMesh: centers, normals scaled by |N|^-1/2 and the
            upper-triangular N N^T weights (NonOrientedSurfaceMesh.cxx:269-308), self norm
            (:312-335), special_product (:364-379), ComputeMatch (:124-165),
            ComputeMatchGradient (:169-260)

The loops follow the reference (face by face, varlin for the varifold gradient).  Pinned by
tests/test_surfaces.py: each match against a brute-force double sum over face pairs, and
each gradient against central differences of its match.
"""
from __future__ import annotations

import numpy as np

from .oracle import ExactKernel


def read_vtk_polydata(path):
    """Points (V x 3) and triangles (F x 3, int) of an ASCII legacy VTK POLYDATA file."""
    tok = open(path).read().split()
    i = tok.index("POINTS")
    nv = int(tok[i + 1])
    pts = np.array(tok[i + 3:i + 3 + 3 * nv], dtype=np.float64).reshape(nv, 3)
    j = tok.index("POLYGONS")
    nf = int(tok[j + 1])
    vals = np.array(tok[j + 3:j + 3 + 4 * nf], dtype=np.int64).reshape(nf, 4)
    assert np.all(vals[:, 0] == 3), "triangles only"
    return pts, vals[:, 1:].copy()


def _conv(Y, W, X, width):
    k = ExactKernel(3, width)
    k.SetSources(Y)
    k.SetWeights(W)
    return k


def current_geometry(V, F):
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    centers = (p0 + p1 + p2) / 3.0
    normals = np.cross(p2 - p0, p1 - p0) / 2
    return centers, normals


def varifold_geometry(V, F):
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    centers = (p0 + p1 + p2) / 3.0
    N = np.cross(p2 - p0, p1 - p0) / 2
    N = N / np.sqrt(np.linalg.norm(N, axis=1, keepdims=True) + 1e-20)
    M = np.stack([N[:, p] * N[:, q] for p in range(3) for q in range(p, 3)], axis=1)
    return centers, N, M


def special_product(M, X):
    """Symmetric matrix from its upper triangle (00, 01, 02, 11, 12, 22) times X, rowwise."""
    S = np.empty(M.shape[:-1] + (3, 3))
    idx = 0
    for p in range(3):
        for q in range(p, 3):
            S[..., p, q] = M[..., idx]
            S[..., q, p] = M[..., idx]
            idx += 1
    return np.einsum("...pq,...q->...p", S, X)


def current_norm2(V, F, width):
    c, n = current_geometry(V, F)
    return float(np.sum(_conv(c, n, c, width).Convolve(c) * n))


def current_match(Vs, Fs, Vt, Ft, width):
    cs, ns = current_geometry(Vs, Fs)
    ct, nt = current_geometry(Vt, Ft)
    sdot = _conv(cs, ns, ct, width).Convolve(ct)
    return current_norm2(Vt, Ft, width) + current_norm2(Vs, Fs, width) - 2.0 * float(np.sum(sdot * nt))


def current_match_gradient(Vs, Fs, Vt, Ft, width):
    cs, ns = current_geometry(Vs, Fs)
    ct, nt = current_geometry(Vt, Ft)
    kS = _conv(cs, ns, cs, width)
    ktS, gS = kS.Convolve(cs), kS.ConvolveGradient(cs, ns)
    kT = _conv(ct, nt, cs, width)
    ktT, gT = kT.Convolve(cs), kT.ConvolveGradient(cs, ns)
    gk = (gS - gT) * 2.0 / 3.0
    grad = np.zeros_like(Vs)
    for f, (i0, i1, i2) in enumerate(Fs):
        e0, e1, e2 = Vs[i2] - Vs[i1], Vs[i0] - Vs[i2], Vs[i1] - Vs[i0]
        kt = ktS[f] - ktT[f]
        grad[i0] += np.cross(e0, kt) + gk[f]
        grad[i1] += np.cross(e1, kt) + gk[f]
        grad[i2] += np.cross(e2, kt) + gk[f]
    return grad


def varifold_norm2(V, F, width):
    c, N, M = varifold_geometry(V, F)
    kw = _conv(c, M, c, width).Convolve(c)
    return float(np.sum(N * special_product(kw, N)))


def varifold_match(Vs, Fs, Vt, Ft, width):
    cs, Ns, Ms = varifold_geometry(Vs, Fs)
    ct, Nt, Mt = varifold_geometry(Vt, Ft)
    sdot = _conv(cs, Ms, ct, width).Convolve(ct)
    cross = float(np.sum(Nt * special_product(sdot, Nt)))
    return varifold_norm2(Vt, Ft, width) + varifold_norm2(Vs, Fs, width) - 2.0 * cross


def varifold_match_gradient(Vs, Fs, Vt, Ft, width):
    cs, Ns, Ms = varifold_geometry(Vs, Fs)
    ct, Nt, Mt = varifold_geometry(Vt, Ft)
    kS = _conv(cs, Ms, cs, width)
    ktS, gS = kS.Convolve(cs), kS.ConvolveGradient(cs)   # (F, 6), (F, 6, 3)
    kT = _conv(ct, Mt, cs, width)
    ktT, gT = kT.Convolve(cs), kT.ConvolveGradient(cs)
    grad = np.zeros_like(Vs)
    for f, (i0, i1, i2) in enumerate(Fs):
        e0, e1, e2 = Vs[i2] - Vs[i1], Vs[i0] - Vs[i2], Vs[i1] - Vs[i0]
        n = Ns[f]
        n2 = float(n @ n)
        if n2 > 1e-20:
            kt = 2.0 * special_product(ktS[f] - ktT[f], n)
            aux = special_product(Ms[f], kt)
            kt = (kt - aux / (2 * n2)) / np.sqrt(n2)
            grad[i0] += np.cross(e0, kt)
            grad[i1] += np.cross(e1, kt)
            grad[i2] += np.cross(e2, kt)
            delta = gS[f] - gT[f]
            gk = np.array([special_product(delta[:, p], n) @ n for p in range(3)]) * (2.0 / 3.0)
            grad[i0] += gk
            grad[i1] += gk
            grad[i2] += gk
    return grad


def brute_force_current(Vs, Fs, Vt, Ft, width):
    cs, ns = current_geometry(Vs, Fs)
    ct, nt = current_geometry(Vt, Ft)

    def ip(c1, n1, c2, n2):
        d2 = np.sum((c1[:, None, :] - c2[None, :, :]) ** 2, axis=2)
        return float(np.sum(np.exp(-d2 / width ** 2) * (n1 @ n2.T)))

    return ip(cs, ns, cs, ns) + ip(ct, nt, ct, nt) - 2 * ip(cs, ns, ct, nt)


def brute_force_varifold(Vs, Fs, Vt, Ft, width):
    def unit_area(V, F):
        p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
        N = np.cross(p2 - p0, p1 - p0) / 2
        return (p0 + p1 + p2) / 3.0, N

    cs, ns = unit_area(Vs, Fs)
    ct, nt = unit_area(Vt, Ft)

    def ip(c1, n1, c2, n2):
        d2 = np.sum((c1[:, None, :] - c2[None, :, :]) ** 2, axis=2)
        a1 = np.linalg.norm(n1, axis=1)
        a2 = np.linalg.norm(n2, axis=1)
        return float(np.sum(np.exp(-d2 / width ** 2) * (n1 @ n2.T) ** 2 / np.outer(a1, a2)))

    return ip(cs, ns, cs, ns) + ip(ct, nt, ct, nt) - 2 * ip(cs, ns, ct, nt)


def brute_force_samples(cs, Ns, ct, Nt, width, kind):
    """Data term between oriented samples (centres, area-weighted normals) by the double sum
    of its definition: current kernel K N_i.N_j, varifold kernel K (N_i.N_j)^2/(|N_i||N_j|)
    with |N| + 1e-20 under the square root as in NonOrientedSurfaceMesh.cxx:269-308."""
    def ip(c1, n1, c2, n2):
        d2 = np.sum((c1[:, None, :] - c2[None, :, :]) ** 2, axis=2)
        K = np.exp(-d2 / width ** 2)
        if kind == "current":
            return float(np.sum(K * (n1 @ n2.T)))
        u1 = n1 / np.sqrt(np.linalg.norm(n1, axis=1, keepdims=True) + 1e-20)
        u2 = n2 / np.sqrt(np.linalg.norm(n2, axis=1, keepdims=True) + 1e-20)
        return float(np.sum(K * (u1 @ u2.T) ** 2))

    return ip(cs, Ns, cs, Ns) + ip(ct, Nt, ct, Nt) - 2 * ip(cs, Ns, ct, Nt)


def samples_varifold_gradient_rows(cs, Ns, ct, Nt, width, rows):
    """d/d centres and d/d (area-weighted) normals of the varifold data term between oriented
    samples, |S|^2 + |T|^2 - 2 <S, T> with <A, B> = sum_ij K(a_i, b_j) (u_i . v_j)^2 and
    u = N / sqrt(|N| + 1e-20) (NonOrientedSurfaceMesh.cxx:269-308), at the source samples
    ``rows`` only, from the kernel sums of NonOrientedSurfaceMesh.cxx:169-260 with every
    sample of both sets as sources: Convolve and ConvolveGradient(X) with the Dv = 6 weights
    M_j = u_j u_j^T (upper triangle), as varifold_match_gradient forms them for a mesh.
      d/d c_i = 2 [u_i^T (dG_S - dG_T)[:, p] u_i]_p
      d/d u_i = 4 (C_S - C_T) u_i,  d/d N_i = (g - u_i (u_i . g) / (2 |u_i|^2)) / |u_i|
    (the chain rule through u = N / sqrt(|N|), as the reference applies it to a triangle's
    normal).  Returns (rows x 3, rows x 3)."""
    def unit_area(N):
        u = N / np.sqrt(np.linalg.norm(N, axis=1, keepdims=True) + 1e-20)
        return u, np.stack([u[:, p] * u[:, q] for p in range(3) for q in range(p, 3)], axis=1)

    us, Ms = unit_area(np.asarray(Ns, dtype=np.float64))
    _, Mt = unit_area(np.asarray(Nt, dtype=np.float64))
    x = np.asarray(cs, dtype=np.float64)[rows]
    u = us[rows]
    kS, kT = _conv(cs, Ms, x, width), _conv(ct, Mt, x, width)
    dC = kS.Convolve(x) - kT.Convolve(x)                    # (r, 6)
    dG = kS.ConvolveGradient(x) - kT.ConvolveGradient(x)    # (r, 6, 3)
    g = 4.0 * special_product(dC, u)
    n2 = np.sum(u * u, axis=1, keepdims=True)
    gN = (g - u * np.sum(u * g, axis=1, keepdims=True) / (2.0 * n2)) / np.sqrt(n2)
    gc = np.stack([2.0 * np.sum(special_product(dG[:, :, p], u) * u, axis=1) for p in range(3)],
                  axis=1)
    return gc, gN

