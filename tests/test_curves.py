"""Oracle polyline data terms (oracle/curves.py, OrientedPolyLine / NonOrientedPolyLine):
matches against brute-force double sums over segment pairs, gradients against central
differences of the matches, in 2D and 3D; the reference's skull polyline as input."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _curves(d, seed=0, n=40, m=33):
    rng = np.random.default_rng(seed)
    s = np.linspace(0, 2 * np.pi, n)
    V = np.zeros((n, d))
    V[:, 0], V[:, 1] = np.cos(s), np.sin(2 * s) * 0.5
    if d == 3:
        V[:, 2] = 0.3 * s / np.pi
    V += rng.normal(scale=0.02, size=V.shape)
    E = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1)
    t = np.linspace(0, 2 * np.pi, m)
    W = np.zeros((m, d))
    W[:, 0], W[:, 1] = 1.1 * np.cos(t) + 0.05, 0.45 * np.sin(2 * t)
    if d == 3:
        W[:, 2] = 0.25 * t / np.pi
    # a closed target with its segments in mixed order and orientation
    F = np.stack([np.arange(m), (np.arange(m) + 1) % m], axis=1)
    F = F[rng.permutation(m)]
    F[::3] = F[::3, ::-1]
    return V, E, W, F


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("kind", ["current", "varifold"])
def test_match_equals_brute_force(d, kind):
    from oracle import curves as oc
    V, E, W, F = _curves(d)
    for width in (0.2, 0.7):
        got = getattr(oc, kind + "_match")(V, E, W, F, width)
        ref = getattr(oc, "brute_force_" + kind)(V, E, W, F, width)
        assert abs(got - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("kind", ["current", "varifold"])
def test_gradient_equals_central_differences(d, kind):
    from oracle import curves as oc
    V, E, W, F = _curves(d, seed=1, n=17, m=13)
    match = getattr(oc, kind + "_match")
    g = getattr(oc, kind + "_match_gradient")(V, E, W, F, 0.5)
    h = 1e-6
    fd = np.zeros_like(V)
    for i in range(V.shape[0]):
        for k in range(d):
            P, M = V.copy(), V.copy()
            P[i, k] += h
            M[i, k] -= h
            fd[i, k] = (match(P, E, W, F, 0.5) - match(M, E, W, F, 0.5)) / (2 * h)
    assert np.linalg.norm(g - fd) <= 1e-6 * np.linalg.norm(fd)


def test_reference_skull_polyline():
    """The 2D skull template of the reference's registration test: 150 points, 144 segments;
    its varifold self-distance is zero and its norm positive."""
    from deformetrica_code_amd.io import read_vtk
    from oracle import curves as oc
    sk = read_vtk(os.path.join(HERE, "golden", "functional", "skulls", "data", "template.vtk"))
    V, E = sk.coordinates(2), sk.segments()
    assert E.shape == (144, 2)
    assert oc.varifold_norm2(V, E, 40.0) > 0
    assert abs(oc.varifold_match(V, E, V, E, 40.0)) <= 1e-9 * oc.varifold_norm2(V, E, 40.0)


def test_segment_incidence():
    """dfm_cell_incidence with two corners: offsets and endpoint ids grouped by point in
    increasing endpoint order; an index outside [0, nv) is refused."""
    from deformetrica_code_amd import _lib
    from deformetrica_code_amd.curves import cell_incidence
    E = np.array([[0, 1], [1, 2], [2, 0], [3, 1]], dtype=np.int32)
    off, inc = cell_incidence(E, 5)
    np.testing.assert_array_equal(off, [0, 2, 5, 7, 8, 8])
    np.testing.assert_array_equal(inc, [0, 5, 1, 2, 7, 3, 4, 6])
    for v in range(5):
        for c in inc[off[v]:off[v + 1]]:
            assert E.reshape(-1)[c] == v
    with pytest.raises(_lib.DfmError, match="outside"):
        cell_incidence(np.array([[0, 5]], dtype=np.int32), 5)
