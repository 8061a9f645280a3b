"""Point-cloud and landmark data terms on the GPU: the PointCloud and Landmark interfaces of
the reference, computed by libdfm_hip.so's dfm_pointcloud_match / dfm_pointcloud_match_gradient
and dfm_landmark_match / dfm_landmark_match_gradient, in dimension 2 or 3.

Mirrors PointCloud.h (SetKernelWidth, GetKernelWidth, GetNumberOfPoints, GetPointCoordinates,
GetPointWeights, GetNormSquared, ComputeMatch, ComputeMatchGradient, DeformedObject) and
Landmark.h (GetNumberOfPoints, GetPointCoordinates, ComputeMatch, ComputeMatchGradient,
DeformedObject) with the reference's error messages (PointCloud.cxx:112-122, 157-169;
Landmark.cxx:290-296, 310-316).  Points and weights stay resident in device memory.

Two behaviours of the reference are kept as written:
- PointCloud::ComputeMatchGradient adds the template-on-template and target-on-template
  gradients (PointCloud.cxx:189): the result is d/dx |S + T|^2, not the derivative of the
  match |S - T|^2;
- the deformed point cloud keeps the template's weights (PointCloud(ex, LP), :68-79).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceArray, Workspace, _sp


def pointcloud_match(sigma, xs, ws, xt, wt, precision="fp64", stream=None, workspace=None):
    """Match (a 1-element DeviceArray) between the template (xs (n, d), ws (n,)) and the
    target (xt (m, d), wt (m,)), all float64 DeviceArrays."""
    L = _lib.load()
    prec, dim = _lib.precision_code(precision), xs.shape[1]
    ws_ = workspace or Workspace()
    wp, wb = ws_.get(L.dfm_pointcloud_workspace(dim, xs.shape[0], xt.shape[0], prec, 0))
    out = DeviceArray((1,))
    _lib.check(L.dfm_pointcloud_match(_sp(stream), dim, float(sigma), prec, xs.ptr, ws.ptr,
                                      xs.shape[0], xt.ptr, wt.ptr, xt.shape[0], out.ptr, wp, wb))
    return out


def pointcloud_match_gradient(sigma, xs, ws, xt, wt, precision="fp64", stream=None,
                              workspace=None, out=None):
    """PointCloud's match gradient (a DeviceArray (n, d)), as the reference writes it."""
    L = _lib.load()
    prec, dim = _lib.precision_code(precision), xs.shape[1]
    ws_ = workspace or Workspace()
    wp, wb = ws_.get(L.dfm_pointcloud_workspace(dim, xs.shape[0], xt.shape[0], prec, 1))
    out = out if out is not None else DeviceArray(xs.shape)
    _lib.check(L.dfm_pointcloud_match_gradient(_sp(stream), dim, float(sigma), prec, xs.ptr,
                                               ws.ptr, xs.shape[0], xt.ptr, wt.ptr, xt.shape[0],
                                               out.ptr, wp, wb))
    return out


def landmark_match(xs, xt, stream=None, workspace=None):
    """|xs - xt|_F^2 (a 1-element DeviceArray) of two float64 DeviceArrays (n, d)."""
    L = _lib.load()
    n, dim = xs.shape
    ws_ = workspace or Workspace()
    wp, wb = ws_.get(L.dfm_landmark_workspace(dim, n))
    out = DeviceArray((1,))
    _lib.check(L.dfm_landmark_match(_sp(stream), dim, xs.ptr, xt.ptr, n, out.ptr, wp, wb))
    return out


def landmark_match_gradient(xs, xt, stream=None, out=None):
    """2 (xs - xt) (a DeviceArray (n, d))."""
    L = _lib.load()
    n, dim = xs.shape
    out = out if out is not None else DeviceArray(xs.shape)
    _lib.check(L.dfm_landmark_match_gradient(_sp(stream), dim, xs.ptr, xt.ptr, n, out.ptr))
    return out


def _points(points):
    points = np.ascontiguousarray(points, dtype=np.float64)
    if points.ndim != 2 or points.shape[1] not in (2, 3):
        raise RuntimeError("point sets are defined in dimension 2 or 3")
    return points


class HIPPointCloud:
    """Point cloud as a sum of weighted Diracs (PointCloud)."""
    TYPE_ERROR = "Abstract Geometries types mismatched"
    WIDTH_ERROR = "Kernel width of point clouds mismatched"

    def __init__(self, points, weights=None, kernel_width=0.0, precision="fp64"):
        points = _points(points)
        n = points.shape[0]
        w = np.ones(n) if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if w.shape != (n,):
            raise RuntimeError("one weight per point")
        self._precision = precision
        self._width = float(kernel_width)
        self._n, self._dim = points.shape
        self._weights = w
        self._X = DeviceArray.from_numpy(points)
        self._W = DeviceArray.from_numpy(w)
        self._ws = Workspace()
        self._stream = None
        self._norm2 = None

    @classmethod
    def from_polydata(cls, polydata, dimension, kernel_width=0.0, precision="fp64"):
        """A point cloud of a VTK file's points, weighted by its point scalars when the
        point data hold one component in all (PointCloud::UpdatePointWeights)."""
        return cls(polydata.coordinates(dimension), polydata.point_weights(), kernel_width,
                   precision)

    def SetKernelWidth(self, h):
        self._width = float(h)
        self._norm2 = None

    def GetKernelWidth(self):
        return self._width

    def GetDimension(self):
        return self._dim

    def GetNumberOfPoints(self):
        return self._n

    def GetPointCoordinates(self):
        return self._X.numpy()

    def SetPointCoordinates(self, points):
        self._X.upload(points)
        self._norm2 = None

    def GetPointWeights(self):
        return self._weights.copy()

    def DeformedObject(self, points):
        """A point cloud with the same weights and kernel width at new point positions
        (PointCloud(ex, LP))."""
        out = type(self).__new__(type(self))
        out.__dict__.update(self.__dict__)
        out._X = DeviceArray.from_numpy(_points(points))
        out._ws = Workspace()
        out._norm2 = None
        return out

    def GetNormSquared(self):
        """|S|^2 = sum_i w_i (K w)(x_i) (UpdateSelfNorm): the match against an empty
        target."""
        if self._norm2 is None:
            ex, ew = DeviceArray((0, self._dim)), DeviceArray((0,))
            m = pointcloud_match(self._width, self._X, self._W, ex, ew, self._precision,
                                 self._stream, self._ws)
            self._norm2 = float(m.numpy()[0])
        return self._norm2

    def _check(self, target):
        if type(target) is not type(self):
            raise RuntimeError(self.TYPE_ERROR)
        if target._width != self._width:
            raise RuntimeError(self.WIDTH_ERROR)
        if target._dim != self._dim:
            raise RuntimeError("point clouds of different dimensions")

    def ComputeMatch(self, target):
        self._check(target)
        m = pointcloud_match(self._width, self._X, self._W, target._X, target._W,
                             self._precision, self._stream, self._ws)
        return float(m.numpy()[0])

    def ComputeMatchGradient(self, target):
        self._check(target)
        g = pointcloud_match_gradient(self._width, self._X, self._W, target._X, target._W,
                                      self._precision, self._stream, self._ws)
        return g.numpy()


class HIPLandmark:
    """Corresponding points (Landmark): no kernel, the match is the squared distance."""

    def __init__(self, points):
        points = _points(points)
        self._n, self._dim = points.shape
        self._X = DeviceArray.from_numpy(points)
        self._ws = Workspace()
        self._stream = None

    @classmethod
    def from_polydata(cls, polydata, dimension):
        return cls(polydata.coordinates(dimension))

    def GetDimension(self):
        return self._dim

    def GetNumberOfPoints(self):
        return self._n

    def GetPointCoordinates(self):
        return self._X.numpy()

    def SetPointCoordinates(self, points):
        self._X.upload(points)

    def DeformedObject(self, points):
        """The landmarks at new point positions (Landmark(ex, LP))."""
        out = type(self).__new__(type(self))
        out.__dict__.update(self.__dict__)
        out._X = DeviceArray.from_numpy(_points(points))
        out._ws = Workspace()
        return out

    def _check(self, target, type_error, count_error):
        if type(target) is not type(self):
            raise RuntimeError(type_error)
        if target._n != self._n:
            raise RuntimeError(count_error)
        if target._dim != self._dim:
            raise RuntimeError("landmarks of different dimensions")

    def ComputeMatch(self, target):
        self._check(target, "Abstract geometries types mismatched",
                    "Landmark object should have the same number of points")
        return float(landmark_match(self._X, target._X, self._stream, self._ws).numpy()[0])

    def ComputeMatchGradient(self, target):
        self._check(target, "Deformable objects types mismatched",
                    "Landmarks object should have the same number of points")
        return landmark_match_gradient(self._X, target._X, self._stream).numpy()


__all__ = ["HIPLandmark", "HIPPointCloud", "landmark_match", "landmark_match_gradient",
           "pointcloud_match", "pointcloud_match_gradient"]
