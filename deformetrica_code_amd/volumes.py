"""Volume-mesh data term on the GPU: the OrientedVolumeMesh interface of the reference
(triangles in 2D, tetrahedra in 3D, each a Dirac at its centre weighted by its volume),
computed by libdfm_hip.so's dfm_volume_match / dfm_volume_match_gradient.

Mirrors OrientedVolumeMesh.h (SetKernelWidth, GetKernelWidth, GetNumberOfCells,
GetNumberOfPoints, GetPointCoordinates, GetNormSquared, ComputeMatch, ComputeMatchGradient,
DeformedObject) with the reference's error messages (OrientedVolumeMesh.cxx:119-127,
160-168, 316-317).  The gradient is the reference's formula as written: the derivative of
the match for positively oriented cells.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .curves import cell_incidence
from .device import DeviceArray, Workspace, _sp


def volume_match(sigma, vs, cs, vt, ct, precision="fp64", stream=None, workspace=None):
    """Match (a 1-element DeviceArray) between template (vs, cs) and target (vt, ct), all
    DeviceArrays (points float64 (n, d), cells int32 (f, d + 1))."""
    L = _lib.load()
    prec, dim = _lib.precision_code(precision), vs.shape[1]
    ws = workspace or Workspace()
    wp, wb = ws.get(L.dfm_volume_workspace(dim, cs.shape[0], ct.shape[0], prec, 0))
    out = DeviceArray((1,))
    _lib.check(L.dfm_volume_match(_sp(stream), dim, float(sigma), prec, vs.ptr, vs.shape[0],
                                  cs.ptr, cs.shape[0], vt.ptr, vt.shape[0], ct.ptr, ct.shape[0],
                                  out.ptr, wp, wb))
    return out


def volume_match_gradient(sigma, vs, cs, offsets, incidences, vt, ct, precision="fp64",
                          stream=None, workspace=None, out=None):
    """d match / d vs (a DeviceArray (n, d)), with the template's incidence list from
    cell_incidence as int32 DeviceArrays."""
    L = _lib.load()
    prec, dim = _lib.precision_code(precision), vs.shape[1]
    ws = workspace or Workspace()
    wp, wb = ws.get(L.dfm_volume_workspace(dim, cs.shape[0], ct.shape[0], prec, 1))
    out = out if out is not None else DeviceArray(vs.shape)
    _lib.check(L.dfm_volume_match_gradient(_sp(stream), dim, float(sigma), prec, vs.ptr,
                                           vs.shape[0], cs.ptr, cs.shape[0], offsets.ptr,
                                           incidences.ptr, vt.ptr, vt.shape[0], ct.ptr,
                                           ct.shape[0], out.ptr, wp, wb))
    return out


class HIPOrientedVolumeMesh:
    """Volume mesh as a measure (OrientedVolumeMesh)."""
    TYPE_ERROR = "Abstract Geometries types mismatched"
    WIDTH_ERROR = "Kernel width of surface currents mismatched"

    def __init__(self, points, cells, kernel_width=0.0, precision="fp64"):
        points = np.ascontiguousarray(points, dtype=np.float64)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        if points.ndim != 2 or points.shape[1] not in (2, 3):
            raise RuntimeError("volume meshes are defined in dimension 2 or 3")
        if cells.ndim != 2 or cells.shape[1] != points.shape[1] + 1:
            raise RuntimeError("Not an admissible cell!")
        self._precision = precision
        self._width = float(kernel_width)
        self._nv, self._dim = points.shape
        self._cells_host = cells
        self._offsets, self._inc = cell_incidence(cells, self._nv)   # checks every index
        self._V = DeviceArray.from_numpy(points)
        self._C = DeviceArray.from_numpy(cells, dtype=np.int32)
        self._dOff = DeviceArray.from_numpy(self._offsets, dtype=np.int32)
        self._dInc = DeviceArray.from_numpy(self._inc, dtype=np.int32)
        self._ws = Workspace()
        self._stream = None
        self._norm2 = None

    def SetKernelWidth(self, h):
        self._width = float(h)
        self._norm2 = None

    def GetKernelWidth(self):
        return self._width

    def GetDimension(self):
        return self._dim

    def GetNumberOfCells(self):
        return self._cells_host.shape[0]

    def GetNumberOfPoints(self):
        return self._nv

    def GetPointCoordinates(self):
        return self._V.numpy()

    def SetPointCoordinates(self, points):
        self._V.upload(points)
        self._norm2 = None

    def DeformedObject(self, points):
        """A volume mesh with the same cells and kernel width at new point positions
        (OrientedVolumeMesh(ex, LP))."""
        out = type(self).__new__(type(self))
        out.__dict__.update(self.__dict__)
        out._V = DeviceArray.from_numpy(points)
        out._ws = Workspace()
        out._norm2 = None
        return out

    def GetNormSquared(self):
        """|S|^2 (UpdateSelfNorm): the match against an empty target."""
        if self._norm2 is None:
            ev = DeviceArray((0, self._dim))
            ec = DeviceArray((0, self._dim + 1), np.int32)
            m = volume_match(self._width, self._V, self._C, ev, ec, self._precision,
                             self._stream, self._ws)
            self._norm2 = float(m.numpy()[0])
        return self._norm2

    def _check(self, target):
        if type(target) is not type(self):
            raise RuntimeError(self.TYPE_ERROR)
        if target._width != self._width:
            raise RuntimeError(self.WIDTH_ERROR)
        if target._dim != self._dim:
            raise RuntimeError("volume meshes of different dimensions")

    def ComputeMatch(self, target):
        self._check(target)
        m = volume_match(self._width, self._V, self._C, target._V, target._C, self._precision,
                         self._stream, self._ws)
        return float(m.numpy()[0])

    def ComputeMatchGradient(self, target):
        self._check(target)
        g = volume_match_gradient(self._width, self._V, self._C, self._dOff, self._dInc,
                                  target._V, target._C, self._precision, self._stream, self._ws)
        return g.numpy()


__all__ = ["HIPOrientedVolumeMesh", "volume_match", "volume_match_gradient"]
