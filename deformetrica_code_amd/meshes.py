"""Synthetic surfaces for the workloads: icospheres, Gaussian-bump displacement fields,
decimated (face-subset) icospheres and oriented sphere samples.

icosphere(level) is the regular icosahedron refined `level` times by edge midpoints pushed
to the sphere: 20 * 4**level triangles and 10 * 4**level + 2 vertices (level 5: 20,480
triangles and 10,242 vertices, the counts of Deformetrica's utilities/meshes/sphere20480.vtk).
Triangles are oriented so that (p2 - p0) x (p1 - p0), the reference's normal
(OrientedSurfaceMesh.cxx:298-326), points inwards; all meshes here share that orientation.
"""
from __future__ import annotations

import numpy as np


def icosphere(level: int, radius: float = 1.0):
    """(vertices (10*4^L+2, 3), faces (20*4^L, 3) int32)."""
    t = (1.0 + 5 ** 0.5) / 2
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    faces = list(F)
    for _ in range(level):
        mid = {}

        def m(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in mid:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                mid[key] = len(verts) - 1
            return mid[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts) * radius, np.array(faces, dtype=np.int32)


def bump_field(points, n_bumps, amplitude, width, rng, radius=1.0):
    """sum_k a_k exp(-|x - c_k|^2 / width^2): c_k uniform on the sphere of `radius`, a_k
    random unit vectors times `amplitude`."""
    c = rng.normal(size=(n_bumps, 3))
    c *= radius / np.linalg.norm(c, axis=1, keepdims=True)
    a = rng.normal(size=(n_bumps, 3))
    a *= amplitude / np.linalg.norm(a, axis=1, keepdims=True)
    d2 = np.sum((points[:, None, :] - c[None, :, :]) ** 2, axis=2)
    return np.exp(-d2 / width ** 2) @ a


_LEVELS = {}


def decimated_icosphere(n_faces, rng, radius=1.0):
    """An icosphere-derived surface with exactly n_faces triangles: the finest icosphere
    with at least that many triangles, keeping a seeded random subset of them (vertices are
    kept whole; unused ones do not enter any data term)."""
    level = 0
    while 20 * 4 ** level < n_faces:
        level += 1
    if (level, radius) not in _LEVELS:
        _LEVELS[(level, radius)] = icosphere(level, radius)
    V, F = _LEVELS[(level, radius)]
    keep = np.sort(rng.choice(F.shape[0], n_faces, replace=False))
    return V.copy(), F[keep]


def sphere_samples(n, rng, log_area_mean=-9.0, log_area_sd=0.3):
    """Oriented samples of the unit sphere: centres uniform on it, normals the unit radial
    direction times a log-normal area."""
    c = rng.normal(size=(n, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    area = np.exp(rng.normal(log_area_mean, log_area_sd, size=(n, 1)))
    return c, c * area
