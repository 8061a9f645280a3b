"""GPU volume-mesh data term (dfm_volume_match / dfm_volume_match_gradient) against the
oracle restatement of OrientedVolumeMesh (oracle/volumes.py), and the reference's own
gradient check (TestVolumeGradient.cxx:79-112) on its meshes, on the GPU."""
import numpy as np
import pytest

from test_volumes import random_mesh, reference_mesh

pytestmark = pytest.mark.gpu
TOL = {"fp64": 1e-10, "fp32c": 1e-6, "fp32x": 1e-6}


def _mesh(V, C, width, prec="fp64"):
    from deformetrica_code_amd.volumes import HIPOrientedVolumeMesh
    return HIPOrientedVolumeMesh(V, C, width, prec)


@pytest.mark.parametrize("prec", ["fp64", "fp32c", "fp32x"])
@pytest.mark.parametrize("case", ["cube-ball", "square-disk", "grid2", "grid3"])
def test_match_and_gradient_match_oracle(case, prec):
    from oracle import volumes as ov
    if case == "cube-ball":
        (V, C), (W, F), width = reference_mesh("VolumeCube.vtk", 3), reference_mesh("VolumeBall.vtk", 3), 1.0
    elif case == "square-disk":
        (V, C), (W, F), width = reference_mesh("VolumeSquare.vtk", 2), reference_mesh("VolumeDisk.vtk", 2), 1.0
    else:
        d = int(case[-1])
        V, C = random_mesh(d, 5, n=12 if d == 3 else 40)
        W, F = random_mesh(d, 6, n=10 if d == 3 else 35)
        W = 1.05 * W + 0.02
        width = 0.15
    print("volume parity", case, prec, flush=True)
    S, T = _mesh(V, C, width, prec), _mesh(W, F, width, prec)
    ref = ov.volume_match(V, C, W, F, width)
    scale = ov.volume_norm2(V, C, width) + ov.volume_norm2(W, F, width)
    assert abs(S.ComputeMatch(T) - ref) <= TOL[prec] * scale
    g_ref = ov.volume_match_gradient(V, C, W, F, width)
    g = S.ComputeMatchGradient(T)
    assert np.linalg.norm(g - g_ref) <= TOL[prec] * np.linalg.norm(g_ref)


@pytest.mark.parametrize("source,target,d", [("VolumeCube.vtk", "VolumeBall.vtk", 3),
                                             ("VolumeSquare.vtk", "VolumeDisk.vtk", 2)])
def test_reference_volume_gradient_check_on_the_gpu(source, target, d):
    """TestVolumeGradient.cxx:79-112 through the GPU object: kernel width 1, forward
    differences of ComputeMatch with h = 1e-5, every component within 1e-5."""
    V, C = reference_mesh(source, d)
    W, F = reference_mesh(target, d)
    S, T = _mesh(V, C, 1.0), _mesh(W, F, 1.0)
    g = S.ComputeMatchGradient(T)
    m0 = S.ComputeMatch(T)
    h = 1e-5
    for i in range(V.shape[0]):
        for k in range(d):
            P = V.copy()
            P[i, k] += h
            assert abs(g[i, k] - (S.DeformedObject(P).ComputeMatch(T) - m0) / h) <= 1e-5


def test_norm_errors_and_bad_cells():
    from deformetrica_code_amd import _lib
    from deformetrica_code_amd.curves import HIPOrientedPolyLine
    from deformetrica_code_amd.device import DeviceArray
    from deformetrica_code_amd.volumes import volume_match
    from oracle import volumes as ov
    V, C = reference_mesh("VolumeSquare.vtk", 2)
    S = _mesh(V, C, 0.5)
    n2 = ov.volume_norm2(V, C, 0.5)
    assert abs(S.GetNormSquared() - n2) <= 1e-12 * n2
    assert abs(S.ComputeMatch(S)) <= 1e-12 * n2
    with pytest.raises(RuntimeError, match="Kernel width of surface currents mismatched"):
        S.ComputeMatch(_mesh(V, C, 0.4))
    with pytest.raises(RuntimeError, match="Abstract Geometries types mismatched"):
        S.ComputeMatch(HIPOrientedPolyLine(V, np.array([[0, 1]]), 0.5))
    bad = C.copy()
    bad[5, 2] = V.shape[0]
    args = [DeviceArray.from_numpy(V), DeviceArray.from_numpy(bad, dtype=np.int32),
            DeviceArray.from_numpy(V), DeviceArray.from_numpy(C, dtype=np.int32)]
    with pytest.raises(_lib.DfmError, match="cell references a vertex index outside"):
        volume_match(0.5, *args)
