"""GPU surface data terms (currents and varifolds) against the oracle restatement of
OrientedSurfaceMesh / NonOrientedSurfaceMesh (oracle/surfaces.py), on the reference's
icosphere meshes and a subdivision of them."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MESH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def _read(name):
    from oracle import surfaces as so
    return so.read_vtk_polydata(os.path.join(MESH, name))


def _subdivide(V, F):
    """Split every triangle into four (edge midpoints pushed to the sphere of radius 0.5)."""
    mid = {}
    verts = [v for v in V]

    def m(a, b):
        key = (min(a, b), max(a, b))
        if key not in mid:
            p = (V[a] + V[b]) / 2
            verts.append(p * (0.5 / np.linalg.norm(p)))
            mid[key] = len(verts) - 1
        return mid[key]

    out = []
    for a, b, c in F:
        ab, bc, ca = m(a, b), m(b, c), m(c, a)
        out += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
    return np.array(verts), np.array(out, dtype=np.int64)


def _pair(name, seed, subdivide=False):
    V, F = _read(name)
    if subdivide:
        V, F = _subdivide(V, F)
    rng = np.random.default_rng(seed)
    bump = 0.05 * np.sin(3 * V[:, [1, 2, 0]])          # smooth template deformation
    Vs = V * 1.05 + bump + rng.normal(scale=0.002, size=V.shape)
    Vt, Ft = _read("sphere80.vtk")
    Vt = Vt + rng.normal(scale=0.01, size=Vt.shape) + 0.03
    return Vs, F, Vt, Ft


def _mesh(kind, V, F, width, prec):
    from deformetrica_code_amd.surfaces import HIPNonOrientedSurfaceMesh, HIPOrientedSurfaceMesh
    cls = HIPOrientedSurfaceMesh if kind == "current" else HIPNonOrientedSurfaceMesh
    return cls(V, F, width, prec)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


TOL = {"fp64": 1e-10, "fp32c": 1e-6, "fp32x": 1e-6}


@pytest.mark.parametrize("prec", ["fp64", "fp32c", "fp32x"])
@pytest.mark.parametrize("kind", ["current", "varifold"])
@pytest.mark.parametrize("mesh,subdivide,width", [("sphere80.vtk", False, 0.3),
                                                  ("sphere1280.vtk", False, 0.1),
                                                  ("sphere1280.vtk", True, 0.05)])
def test_match_and_gradient_match_oracle(kind, prec, mesh, subdivide, width):
    """Match to TOL relative to |S|^2 + |T|^2 (the match is a difference of such terms);
    gradient to TOL norm-wise."""
    from oracle import surfaces as so
    Vs, Fs, Vt, Ft = _pair(mesh, 3, subdivide)
    S, T = _mesh(kind, Vs, Fs, width, prec), _mesh(kind, Vt, Ft, width, prec)
    ref = getattr(so, kind + "_match")(Vs, Fs, Vt, Ft, width)
    scale = getattr(so, kind + "_norm2")(Vs, Fs, width) + getattr(so, kind + "_norm2")(Vt, Ft, width)
    assert abs(S.ComputeMatch(T) - ref) <= TOL[prec] * scale
    g_ref = getattr(so, kind + "_match_gradient")(Vs, Fs, Vt, Ft, width)
    assert rel(S.ComputeMatchGradient(T), g_ref) <= TOL[prec]


@pytest.mark.parametrize("kind", ["current", "varifold"])
def test_norm_squared_and_self_match(kind):
    from oracle import surfaces as so
    Vs, Fs, _, _ = _pair("sphere1280.vtk", 4)
    S = _mesh(kind, Vs, Fs, 0.1, "fp64")
    n2 = getattr(so, kind + "_norm2")(Vs, Fs, 0.1)
    assert abs(S.GetNormSquared() - n2) <= 1e-12 * n2
    # the match of a surface with itself is 0 up to rounding, its gradient exactly 0
    assert abs(S.ComputeMatch(S)) <= 1e-12 * n2
    np.testing.assert_array_equal(S.ComputeMatchGradient(S), np.zeros_like(Vs))


@pytest.mark.parametrize("kind", ["current", "varifold"])
def test_gpu_gradient_matches_central_differences(kind):
    Vs, Fs, Vt, Ft = _pair("sphere80.vtk", 5)
    w = 0.25
    S, T = _mesh(kind, Vs, Fs, w, "fp64"), _mesh(kind, Vt, Ft, w, "fp64")
    g = S.ComputeMatchGradient(T)
    rng = np.random.default_rng(6)
    eps = 1e-6
    for i in rng.choice(Vs.shape[0], 8, replace=False):
        for k in range(3):
            p, m = Vs.copy(), Vs.copy()
            p[i, k] += eps
            m[i, k] -= eps
            fd = (S.DeformedObject(p).ComputeMatch(T) - S.DeformedObject(m).ComputeMatch(T)) / (2 * eps)
            assert abs(g[i, k] - fd) <= 1e-6 * np.abs(g).max()


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_gradient_is_deterministic(prec):
    Vs, Fs, Vt, Ft = _pair("sphere1280.vtk", 7, True)
    S, T = _mesh("varifold", Vs, Fs, 0.05, prec), _mesh("varifold", Vt, Ft, 0.05, prec)
    np.testing.assert_array_equal(S.ComputeMatchGradient(T), S.ComputeMatchGradient(T))
    assert S.ComputeMatch(T) == S.ComputeMatch(T)


def test_reference_errors():
    Vs, Fs, Vt, Ft = _pair("sphere80.vtk", 8)
    S = _mesh("current", Vs, Fs, 0.3, "fp64")
    with pytest.raises(RuntimeError, match="Kernel width of surface currents mismatched"):
        S.ComputeMatch(_mesh("current", Vt, Ft, 0.2, "fp64"))
    with pytest.raises(RuntimeError, match="AbstractGeometries types mismatched"):
        S.ComputeMatch(_mesh("varifold", Vt, Ft, 0.3, "fp64"))
    U = _mesh("varifold", Vs, Fs, 0.3, "fp64")
    with pytest.raises(RuntimeError, match="Kernel width of surface unsigned currents mismatched"):
        U.ComputeMatch(_mesh("varifold", Vt, Ft, 0.2, "fp64"))


def test_out_of_range_face_is_an_error_on_the_device_path():
    """The geometry kernel bounds-checks every index and the call reports it."""
    from deformetrica_code_amd import _lib
    from deformetrica_code_amd.device import DeviceArray
    from deformetrica_code_amd.surfaces import surface_match
    Vs, Fs, Vt, Ft = _pair("sphere80.vtk", 9)
    bad = Fs.copy()
    bad[17, 1] = Vs.shape[0]
    args = [DeviceArray.from_numpy(Vs), DeviceArray.from_numpy(bad, dtype=np.int32),
            DeviceArray.from_numpy(Vt), DeviceArray.from_numpy(Ft, dtype=np.int32)]
    with pytest.raises(_lib.DfmError, match="outside"):
        surface_match("current", 0.3, *args)


def _samples(seed, ns, nt):
    from deformetrica_code_amd.meshes import sphere_samples
    rng = np.random.default_rng(seed)
    cs, Ns = sphere_samples(ns, rng, -3.0, 0.3)
    ct, Nt = sphere_samples(nt, rng, -3.0, 0.3)
    return cs, Ns, ct + 0.02, Nt


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
@pytest.mark.parametrize("kind", ["current", "varifold"])
def test_samples_match_and_gradient(kind, prec):
    """Oriented-sample data term against the oracle's double sum; d/d centres and
    d/d normals against central differences of that double sum."""
    from deformetrica_code_amd.device import DeviceArray
    from deformetrica_code_amd.surfaces import samples_match_gradient
    from oracle import surfaces as so
    cs, Ns, ct, Nt = _samples(11, 700, 500)
    w = 0.15
    m, gc, gn = samples_match_gradient(kind, w, *(DeviceArray.from_numpy(v) for v in (cs, Ns, ct, Nt)),
                                       precision=prec)
    ref = so.brute_force_samples(cs, Ns, ct, Nt, w, kind)
    scale = so.brute_force_samples(cs, Ns, ct[:0], Nt[:0], w, kind) + \
        so.brute_force_samples(ct, Nt, cs[:0], Ns[:0], w, kind)
    tol = 1e-10 if prec == "fp64" else 1e-6
    assert abs(m.numpy()[0] - ref) <= tol * scale
    gc, gn = gc.numpy(), gn.numpy()
    rng = np.random.default_rng(12)
    eps = 1e-6
    for i in rng.choice(cs.shape[0], 5, replace=False):
        for k in range(3):
            for arr, g in ((cs, gc), (Ns, gn)):
                p, q = arr.copy(), arr.copy()
                p[i, k] += eps * (1 if arr is cs else abs(Ns[i]).max())
                q[i, k] -= eps * (1 if arr is cs else abs(Ns[i]).max())
                h = p[i, k] - q[i, k]
                args_p = (p, Ns) if arr is cs else (cs, p)
                args_q = (q, Ns) if arr is cs else (cs, q)
                fd = (so.brute_force_samples(*args_p, ct, Nt, w, kind) -
                      so.brute_force_samples(*args_q, ct, Nt, w, kind)) / h
                assert abs(g[i, k] - fd) <= (1e-6 if prec == "fp64" else 1e-4) * np.abs(g).max()
