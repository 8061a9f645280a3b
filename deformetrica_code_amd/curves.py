"""Polyline data terms on the GPU: the OrientedPolyLine (currents) and NonOrientedPolyLine
(varifolds) interface of the reference, computed by libdfm_hip.so's dfm_curve_match /
dfm_curve_match_gradient, in dimension 2 or 3.

Mirrors OrientedPolyLine.h and NonOrientedPolyLine.h (SetKernelWidth, GetKernelWidth,
GetNumberOfCells, GetNumberOfPoints, GetPointCoordinates, GetNormSquared, ComputeMatch,
ComputeMatchGradient, DeformedObject) with the reference's error messages
(OrientedPolyLine.cxx:134-143, 179-186, 263-264; NonOrientedPolyLine.cxx:131-140, 179-186).
Points and segments stay resident in device memory; the point->endpoint incidence list is
built once per polyline.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray, Workspace, _sp

_KIND = {"current": _lib.CURRENT, "varifold": _lib.VARIFOLD}


def cell_incidence(cells, nv):
    """Point -> corner incidence list (offsets (nv + 1), incidences (c * ncells)) of a cell
    list with c points per cell, as built by dfm_cell_incidence."""
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    n, c = cells.shape
    offsets = np.zeros(nv + 1, dtype=np.int32)
    inc = np.zeros(max(c * n, 1), dtype=np.int32)
    _lib.check(_lib.load().dfm_cell_incidence(cells.ctypes.data_as(ctypes.c_void_p), n, c, nv,
                                              offsets.ctypes.data_as(ctypes.c_void_p),
                                              inc.ctypes.data_as(ctypes.c_void_p)))
    return offsets, inc[:c * n]


def curve_match(kind, sigma, vs, es, vt, et, precision="fp64", stream=None, workspace=None):
    """Match (a 1-element DeviceArray) between template (vs, es) and target (vt, et), all
    DeviceArrays (points float64 (n, d), segments int32 (e, 2))."""
    L = _lib.load()
    k, prec, dim = _KIND[kind], _lib.precision_code(precision), vs.shape[1]
    ws = workspace or Workspace()
    wp, wb = ws.get(L.dfm_curve_workspace(k, dim, es.shape[0], et.shape[0], prec, 0))
    out = DeviceArray((1,))
    _lib.check(L.dfm_curve_match(_sp(stream), k, dim, float(sigma), prec, vs.ptr, vs.shape[0],
                                 es.ptr, es.shape[0], vt.ptr, vt.shape[0], et.ptr, et.shape[0],
                                 out.ptr, wp, wb))
    return out


def curve_match_gradient(kind, sigma, vs, es, offsets, incidences, vt, et, precision="fp64",
                         stream=None, workspace=None, out=None):
    """d match / d vs (a DeviceArray (n, d)); offsets/incidences are the template's
    incidence list from cell_incidence, as int32 DeviceArrays."""
    L = _lib.load()
    k, prec, dim = _KIND[kind], _lib.precision_code(precision), vs.shape[1]
    ws = workspace or Workspace()
    wp, wb = ws.get(L.dfm_curve_workspace(k, dim, es.shape[0], et.shape[0], prec, 1))
    out = out if out is not None else DeviceArray(vs.shape)
    _lib.check(L.dfm_curve_match_gradient(_sp(stream), k, dim, float(sigma), prec, vs.ptr,
                                          vs.shape[0], es.ptr, es.shape[0], offsets.ptr,
                                          incidences.ptr, vt.ptr, vt.shape[0], et.ptr,
                                          et.shape[0], out.ptr, wp, wb))
    return out


class _PolyLine:
    KIND = None
    TYPE_ERROR = "Deformable object types mismatched"
    WIDTH_ERROR = "Kernel width of curve currents mismatched"
    GRAD_TYPE_ERROR = "Deformable objects types mismatched"
    GRAD_WIDTH_ERROR = "Kernel width of curve currents mismatched"

    def __init__(self, points, segments, kernel_width=0.0, precision="fp64"):
        points = np.ascontiguousarray(points, dtype=np.float64)
        segments = np.ascontiguousarray(segments, dtype=np.int32)
        if points.ndim != 2 or points.shape[1] not in (2, 3):
            raise RuntimeError("polylines are defined in dimension 2 or 3")
        if segments.ndim != 2 or segments.shape[1] != 2:
            raise RuntimeError("Not a polygonal line!")
        self._precision = precision
        self._width = float(kernel_width)
        self._nv, self._dim = points.shape
        self._segments_host = segments
        self._offsets, self._inc = cell_incidence(segments, self._nv)   # checks every index
        self._V = DeviceArray.from_numpy(points)
        self._E = DeviceArray.from_numpy(segments, dtype=np.int32)
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

    def GetDimension(self):
        return self._dim

    def GetNumberOfCells(self):
        return self._segments_host.shape[0]

    def GetNumberOfPoints(self):
        return self._nv

    def GetPointCoordinates(self):
        return self._V.numpy()

    def SetPointCoordinates(self, points):
        self._V.upload(points)
        self._norm2 = None

    def DeformedObject(self, points):
        """A polyline with the same segments and kernel width at new point positions
        (OrientedPolyLine(ex, LP), OrientedPolyLine.cxx:75-87)."""
        out = type(self).__new__(type(self))
        out.__dict__.update(self.__dict__)
        out._V = DeviceArray.from_numpy(points)
        out._ws = Workspace()
        out._norm2 = None
        return out

    def GetNormSquared(self):
        """|S|^2 (UpdateSelfNorm): the match against an empty target."""
        if self._norm2 is None:
            ev, ee = DeviceArray((0, self._dim)), DeviceArray((0, 2), np.int32)
            m = curve_match(self.KIND, self._width, self._V, self._E, ev, ee, self._precision,
                            self._stream, self._ws)
            self._norm2 = float(m.numpy()[0])
        return self._norm2

    def _check(self, target, type_error, width_error):
        if type(target) is not type(self):
            raise RuntimeError(type_error)
        if target._width != self._width:
            raise RuntimeError(width_error)
        if target._dim != self._dim:
            raise RuntimeError("polylines of different dimensions")

    def ComputeMatch(self, target):
        self._check(target, self.TYPE_ERROR, self.WIDTH_ERROR)
        m = curve_match(self.KIND, self._width, self._V, self._E, target._V, target._E,
                        self._precision, self._stream, self._ws)
        return float(m.numpy()[0])

    def ComputeMatchGradient(self, target):
        self._check(target, self.GRAD_TYPE_ERROR, self.GRAD_WIDTH_ERROR)
        g = curve_match_gradient(self.KIND, self._width, self._V, self._E, self._dOff,
                                 self._dInc, target._V, target._E, self._precision,
                                 self._stream, self._ws)
        return g.numpy()


class HIPOrientedPolyLine(_PolyLine):
    """Polyline as a current (OrientedPolyLine)."""
    KIND = "current"


class HIPNonOrientedPolyLine(_PolyLine):
    """Polyline as a varifold (NonOrientedPolyLine)."""
    KIND = "varifold"
    GRAD_WIDTH_ERROR = "Kernel width of surface currents mismatched"


__all__ = ["HIPOrientedPolyLine", "HIPNonOrientedPolyLine", "cell_incidence", "curve_match",
           "curve_match_gradient"]
