"""Oracle volume-mesh data term (oracle/volumes.py, OrientedVolumeMesh): the match against a
brute-force double sum over cells, the gradient against central differences on positively
oriented meshes, and the reference's own gradient check (TestVolumeGradient.cxx:79-112:
forward differences with h = 1e-5 within 1e-5, kernel width 1) on its volume meshes."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
VOL = os.path.join(HERE, "golden", "volumes")


def positive(V, C):
    """Swap two corners of every negatively oriented cell."""
    C = np.array(C)
    P = V[C]
    neg = np.linalg.det(P[:, 1:] - P[:, :1]) < 0
    C[neg, 0], C[neg, 1] = C[neg, 1].copy(), C[neg, 0].copy()
    return C


def random_mesh(d, seed, n=7):
    """Jittered grid cells: triangles (2D) or the six tetrahedra of each cube (3D)."""
    rng = np.random.default_rng(seed)
    g = np.arange(n)
    if d == 2:
        V = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2).astype(float)
        idx = lambda i, j: i * n + j
        C = []
        for i in range(n - 1):
            for j in range(n - 1):
                C += [(idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)),
                      (idx(i, j), idx(i + 1, j + 1), idx(i, j + 1))]
    else:
        V = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(float)
        idx = lambda i, j, k: (i * n + j) * n + k
        C = []
        for i in range(n - 1):
            for j in range(n - 1):
                for k in range(n - 1):
                    c = [idx(i + a, j + b, k + e) for a in (0, 1) for b in (0, 1) for e in (0, 1)]
                    # the six tetrahedra around the main diagonal c0 - c7
                    for p, q in ((1, 3), (3, 2), (2, 6), (6, 4), (4, 5), (5, 1)):
                        C.append((c[0], c[p], c[q], c[7]))
    V = V / (n - 1) + rng.normal(scale=0.03, size=V.shape)
    return V, positive(V, C)


def reference_mesh(name, d):
    from deformetrica_code_amd.io import read_vtk
    m = read_vtk(os.path.join(VOL, name))
    return m.coordinates(d), m.volume_cells(d)


@pytest.mark.parametrize("d", [2, 3])
def test_match_equals_brute_force(d):
    from oracle import volumes as ov
    V, C = random_mesh(d, 1)
    W, F = random_mesh(d, 2)
    W = 1.1 * W + 0.05
    for width in (0.2, 0.6):
        got = ov.volume_match(V, C, W, F, width)
        ref = ov.brute_force_volume(V, C, W, F, width)
        assert abs(got - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("d", [2, 3])
def test_gradient_equals_central_differences(d):
    """Up to the reference's float(1/3) in 3D (relative 3e-8 below 1/3), the gradient is the
    derivative of the match for positively oriented cells."""
    from oracle import volumes as ov
    V, C = random_mesh(d, 3, n=4)
    W, F = random_mesh(d, 4, n=5)
    W = 0.9 * W + 0.1
    g = ov.volume_match_gradient(V, C, W, F, 0.5)
    h = 1e-6
    fd = np.zeros_like(V)
    for i in range(V.shape[0]):
        for k in range(d):
            P, M = V.copy(), V.copy()
            P[i, k] += h
            M[i, k] -= h
            fd[i, k] = (ov.volume_match(P, C, W, F, 0.5) - ov.volume_match(M, C, W, F, 0.5)) / (2 * h)
    assert np.linalg.norm(g - fd) <= 1e-6 * np.linalg.norm(fd)


@pytest.mark.parametrize("source,target,d", [("VolumeCube.vtk", "VolumeBall.vtk", 3),
                                             ("VolumeSquare.vtk", "VolumeDisk.vtk", 2)])
def test_reference_volume_gradient_check(source, target, d):
    """TestVolumeGradient.cxx:79-112 on the oracle: kernel width 1, forward differences with
    h = 1e-5, every component within 1e-5 absolute."""
    from oracle import volumes as ov
    V, C = reference_mesh(source, d)
    W, F = reference_mesh(target, d)
    assert (np.linalg.det(V[C][:, 1:] - V[C][:, :1]) > 0).all()
    g = ov.volume_match_gradient(V, C, W, F, 1.0)
    m0 = ov.volume_match(V, C, W, F, 1.0)
    h = 1e-5
    for i in range(V.shape[0]):
        for k in range(d):
            P = V.copy()
            P[i, k] += h
            assert abs(g[i, k] - (ov.volume_match(P, C, W, F, 1.0) - m0) / h) <= 1e-5


def test_volume_cells_of_the_reference_files():
    V, C = reference_mesh("VolumeCube.vtk", 3)
    assert V.shape == (30, 3) and C.shape == (111, 4)
    V, C = reference_mesh("VolumeDisk.vtk", 2)
    assert V.shape == (100, 2) and C.shape == (178, 3)
