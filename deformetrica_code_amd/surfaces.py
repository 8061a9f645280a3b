"""Surface data terms on the GPU: the OrientedSurfaceMesh (currents) and
NonOrientedSurfaceMesh (varifolds) interface of the reference, computed by
libdfm_hip.so's dfm_surface_match / dfm_surface_match_gradient.

Mirrors OrientedSurfaceMesh.h:83-145 and NonOrientedSurfaceMesh.h (SetKernelWidth,
GetKernelWidth, GetNumberOfCells, GetNumberOfPoints, GetPointCoordinates, GetNormSquared,
ComputeMatch, ComputeMatchGradient, DeformedObject) with the reference's error messages
(OrientedSurfaceMesh.cxx:130-137, NonOrientedSurfaceMesh.cxx:129-136, 174-181).
Vertices and triangles stay resident in device memory; the vertex->corner incidence
list of the triangles is built once per mesh.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray, Stream, Workspace, _sp

_KIND = {"current": _lib.CURRENT, "varifold": _lib.VARIFOLD}


def mesh_incidence(faces, nv):
    """Vertex -> corner incidence list (offsets (nv + 1), incidences (3 nf)) of a mesh."""
    faces = np.ascontiguousarray(faces, dtype=np.int32)
    nf = faces.shape[0]
    offsets = np.zeros(nv + 1, dtype=np.int32)
    inc = np.zeros(max(3 * nf, 1), dtype=np.int32)
    _lib.check(_lib.load().dfm_mesh_incidence(faces.ctypes.data_as(ctypes.c_void_p), nf, nv,
                                              offsets.ctypes.data_as(ctypes.c_void_p),
                                              inc.ctypes.data_as(ctypes.c_void_p)))
    return offsets, inc[:3 * nf]


def surface_match(kind, sigma, vs, fs, vt, ft, precision="fp64", stream=None, workspace=None):
    """Match (a 1-element DeviceArray) between template (vs, fs) and target (vt, ft), all
    DeviceArrays (vertices float64 (n, 3), faces int32 (f, 3))."""
    L = _lib.load()
    k, prec = _KIND[kind], _lib.precision_code(precision)
    ws = workspace or Workspace()
    wp, wb = ws.get(L.dfm_surface_workspace(k, fs.shape[0], ft.shape[0], prec, 0))
    out = DeviceArray((1,))
    _lib.check(L.dfm_surface_match(_sp(stream), k, float(sigma), prec, vs.ptr, vs.shape[0], fs.ptr,
                                   fs.shape[0], vt.ptr, vt.shape[0], ft.ptr, ft.shape[0], out.ptr,
                                   wp, wb))
    return out


def surface_match_gradient(kind, sigma, vs, fs, offsets, incidences, vt, ft, precision="fp64",
                           stream=None, workspace=None, out=None):
    """d match / d vs (a DeviceArray (n, 3)); offsets/incidences are the template's
    incidence list from mesh_incidence, as int32 DeviceArrays."""
    L = _lib.load()
    k, prec = _KIND[kind], _lib.precision_code(precision)
    ws = workspace or Workspace()
    wp, wb = ws.get(L.dfm_surface_workspace(k, fs.shape[0], ft.shape[0], prec, 1))
    out = out if out is not None else DeviceArray(vs.shape)
    _lib.check(L.dfm_surface_match_gradient(_sp(stream), k, float(sigma), prec, vs.ptr, vs.shape[0],
                                            fs.ptr, fs.shape[0], offsets.ptr, incidences.ptr,
                                            vt.ptr, vt.shape[0], ft.ptr, ft.shape[0], out.ptr, wp, wb))
    return out


def samples_match_gradient(kind, sigma, cs, Ns, ct, Nt, precision="fp32c", gradient=True,
                           stream=None, workspace=None):
    """Data term between oriented samples (DeviceArrays of centres and area-weighted
    normals): (match (1,), d/d cs, d/d Ns) as DeviceArrays (gradients None if not asked)."""
    L = _lib.load()
    k, prec = _KIND[kind], _lib.precision_code(precision)
    ws = workspace or Workspace()
    ns, nt = cs.shape[0], ct.shape[0]
    wp, wb = ws.get(L.dfm_samples_workspace(k, ns, nt, prec, int(gradient)))
    out = DeviceArray((1,))
    gc = DeviceArray((ns, 3)) if gradient else None
    gn = DeviceArray((ns, 3)) if gradient else None
    _lib.check(L.dfm_samples_match_gradient(_sp(stream), k, float(sigma), prec, cs.ptr, Ns.ptr, ns,
                                            ct.ptr, Nt.ptr, nt, out.ptr, gc.ptr if gc else None,
                                            gn.ptr if gn else None, wp, wb))
    return out, gc, gn


def reorient_normals(vertices, faces):
    """The triangle order OrientedSurfaceMesh gets with ``reorient-normals`` on
    (OrientedSurfaceMesh.cxx:274-278: vtkPolyDataNormals with ConsistencyOn,
    AutoOrientNormalsOn, FlipNormalsOn; myvtkPolyDataNormals.cxx:212-305, 549-...):

    per connected component, the seed is the unvisited triangle at the leftmost point whose
    normal (p1 - p0) x (p2 - p0) has the largest |x| component; it is oriented so that this
    normal points to +x (outward-left, then flipped), and every triangle reached through a
    shared edge is reversed where it runs that edge in the same direction as its visited
    neighbour.  For a closed surface the result has (p1 - p0) x (p2 - p0) pointing inward,
    so the current's normals (p2 - p0) x (p1 - p0) / 2 point outward.  Reversing a
    triangle reverses its vertex order, as vtkPolyData::ReverseCell does.  Returns a new
    (F, 3) int32 array."""
    V = np.asarray(vertices, dtype=np.float64)
    F = np.array(faces, dtype=np.int64, copy=True)
    nf = F.shape[0]
    if nf == 0:
        return F.astype(np.int32)
    edges = {}
    cells_at = [[] for _ in range(V.shape[0])]
    for c, tri in enumerate(F):
        for k in range(3):
            a, b = int(tri[k]), int(tri[(k + 1) % 3])
            edges.setdefault((min(a, b), max(a, b)), []).append(c)
            cells_at[a].append(c)
    visited = np.zeros(nf, dtype=bool)

    def normal_x(c):
        p0, p1, p2 = V[F[c, 0]], V[F[c, 1]], V[F[c, 2]]
        return np.cross(p1 - p0, p2 - p0)[0]

    def runs(c, a, b):  # does triangle c traverse a -> b?
        t = F[c]
        return any(int(t[k]) == a and int(t[(k + 1) % 3]) == b for k in range(3))

    for pt in np.argsort(V[:, 0], kind="stable"):
        best, best_abs = -1, 0.0
        for c in cells_at[pt]:
            if visited[c]:
                continue
            nx = normal_x(c)
            if abs(nx) > best_abs:
                best, best_abs = c, abs(nx)
        if best < 0:
            continue
        if not normal_x(best) > 0:  # reverse flag XOR flip: make the seed's normal face +x
            F[best] = F[best][::-1]
        visited[best] = True
        wave = [best]
        while wave:  # TraverseAndOrder
            nxt = []
            for c in wave:
                t = F[c]
                for k in range(3):
                    a, b = int(t[k]), int(t[(k + 1) % 3])
                    for nb in edges[(min(a, b), max(a, b))]:
                        if nb == c or visited[nb]:
                            continue
                        if runs(nb, a, b):
                            F[nb] = F[nb][::-1]
                        visited[nb] = True
                        nxt.append(nb)
            wave = nxt
    return F.astype(np.int32)


class _SurfaceMesh:
    KIND = None
    TYPE_ERROR = "AbstractGeometries types mismatched"
    WIDTH_ERROR = "Kernel width of surface currents mismatched"
    GRAD_TYPE_ERROR = TYPE_ERROR
    GRAD_WIDTH_ERROR = WIDTH_ERROR

    def __init__(self, vertices, faces, kernel_width=0.0, precision="fp64"):
        vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        if vertices.ndim != 2 or vertices.shape[1] != 3:
            raise RuntimeError("Cannot create Surface Mesh Object in Dimension other than 3")
        if faces.ndim != 2 or faces.shape[1] != 3:
            raise RuntimeError("Not a triangle cell!")
        self._precision = precision
        self._width = float(kernel_width)
        self._nv = vertices.shape[0]
        self._faces_host = faces
        self._offsets, self._inc = mesh_incidence(faces, self._nv)   # checks every index
        self._V = DeviceArray.from_numpy(vertices)
        self._F = DeviceArray.from_numpy(faces, dtype=np.int32)
        self._dOff = DeviceArray.from_numpy(self._offsets, dtype=np.int32)
        self._dInc = DeviceArray.from_numpy(self._inc, dtype=np.int32)
        self._ws = Workspace()
        self._stream = None
        self._norm2 = None

    # --- reference accessors ---
    def SetKernelWidth(self, h):
        self._width = float(h)
        self._norm2 = None

    def GetKernelWidth(self):
        return self._width

    def GetNumberOfCells(self):
        return self._faces_host.shape[0]

    def GetNumberOfPoints(self):
        return self._nv

    def GetPointCoordinates(self):
        return self._V.numpy()

    def SetPointCoordinates(self, vertices):
        self._V.upload(vertices)
        self._norm2 = None

    def DeformedObject(self, vertices):
        """A mesh with the same triangles and kernel width at new vertex positions
        (doDeformedObject, OrientedSurfaceMesh.h:156-158)."""
        out = type(self).__new__(type(self))
        out.__dict__.update(self.__dict__)
        out._V = DeviceArray.from_numpy(vertices)
        out._ws = Workspace()
        out._norm2 = None
        return out

    def GetNormSquared(self):
        """|S|^2 (UpdateSelfNorm): the match against an empty target."""
        if self._norm2 is None:
            empty_v, empty_f = DeviceArray((0, 3)), DeviceArray((0, 3), np.int32)
            m = surface_match(self.KIND, self._width, self._V, self._F, empty_v, empty_f,
                              self._precision, self._stream, self._ws)
            self._norm2 = float(m.numpy()[0])
        return self._norm2

    def _check(self, target, type_error, width_error):
        if type(target) is not type(self):
            raise RuntimeError(type_error)
        if target._width != self._width:
            raise RuntimeError(width_error)

    def ComputeMatch(self, target):
        self._check(target, self.TYPE_ERROR, self.WIDTH_ERROR)
        m = surface_match(self.KIND, self._width, self._V, self._F, target._V, target._F,
                          self._precision, self._stream, self._ws)
        return float(m.numpy()[0])

    def ComputeMatchGradient(self, target):
        self._check(target, self.GRAD_TYPE_ERROR, self.GRAD_WIDTH_ERROR)
        g = surface_match_gradient(self.KIND, self._width, self._V, self._F, self._dOff,
                                   self._dInc, target._V, target._F, self._precision,
                                   self._stream, self._ws)
        return g.numpy()


class HIPOrientedSurfaceMesh(_SurfaceMesh):
    """Surface as a current (OrientedSurfaceMesh)."""
    KIND = "current"


class HIPNonOrientedSurfaceMesh(_SurfaceMesh):
    """Surface as a varifold (NonOrientedSurfaceMesh)."""
    KIND = "varifold"
    TYPE_ERROR = "Deformable object types mismatched"
    WIDTH_ERROR = "Kernel width of surface unsigned currents mismatched"
    GRAD_TYPE_ERROR = "Deformable objects types mismatched"
    GRAD_WIDTH_ERROR = "Kernel width of surface currents mismatched"


__all__ = ["HIPOrientedSurfaceMesh", "HIPNonOrientedSurfaceMesh", "mesh_incidence",
           "surface_match", "surface_match_gradient", "samples_match_gradient", "Stream"]
