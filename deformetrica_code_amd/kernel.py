"""Host-side mirror of the reference kernel plugin interface.

``HIPExactKernel`` has the method set, argument meaning and error behaviour of
Deformetrica's ``CUDAExactKernel<double, D>`` (src/support/kernels/CUDAExactKernel.h:29-81),
the exact-kernel type it replaces: the kernel owns copies of its sources and weights
(AbstractKernel.h:93-108), clamps the width at 1e-10 (AbstractKernel.h:85-91), and the
GPU ops are Convolve, the three ConvolveGradient overloads, the three ConvolveHessian
overloads (which CUDAExactKernel inherits from ExactKernel on the CPU) and
ConvolveSpecialHessian.
Matrices are numpy float64 arrays (rows x cols).  Every op runs on the gfx950 library
through its synchronous host-pointer entry points; nothing is computed in Python.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_D = ctypes.POINTER(ctypes.c_double)


def _arr(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(_D)


class HIPExactKernel:
    """Gaussian kernel K(x, y) = exp(-|x - y|^2 / h^2) evaluated on the GPU."""

    def __init__(self, dim: int, kernel_width: float = 1.0, precision: str = "fp64"):
        if dim not in (2, 3):
            raise ValueError("PointDim must be 2 or 3")
        self.dim = dim
        self.precision = _lib.precision_code(precision)
        self.SetKernelWidth(kernel_width)
        self._Y = np.zeros((0, dim))
        self._W = np.zeros((0, dim))
        self._modified = True

    def Clone(self) -> "HIPExactKernel":
        k = HIPExactKernel(self.dim, self._h, self.precision)
        k._Y, k._W, k._modified = self._Y.copy(), self._W.copy(), self._modified
        return k

    # AbstractKernel.h:82-115
    def GetKernelWidth(self) -> float:
        return self._h

    def SetKernelWidth(self, h: float) -> None:
        self._h = max(float(h), 1e-10)
        self._modified = True

    def GetSources(self) -> np.ndarray:
        return self._Y

    def SetSources(self, Y) -> None:
        Y = _arr(Y)
        if Y.ndim != 2 or Y.shape[1] != self.dim:
            raise ValueError("sources must be (N, %d)" % self.dim)
        self._Y = Y.copy()
        self._modified = True

    def GetWeights(self) -> np.ndarray:
        return self._W

    def SetWeights(self, W) -> None:
        W = _arr(W)
        if W.ndim != 2:
            raise ValueError("weights must be a matrix")
        self._W = W.copy()
        self._modified = True

    def IsModified(self) -> bool:
        return self._modified

    def _check_rows(self):
        if self._Y.shape[0] != self._W.shape[0]:
            raise RuntimeError("Sources and weights count mismatch")

    def _targets(self, X) -> np.ndarray:
        X = _arr(X)
        if X.ndim != 2 or X.shape[1] != self.dim:
            raise ValueError("points must be (N, %d)" % self.dim)
        return X

    # CUDAExactKernel.cxx:41-90
    def Convolve(self, X) -> np.ndarray:
        self._check_rows()
        X = self._targets(X)
        dv = self._W.shape[1]
        out = np.empty((X.shape[0], dv))
        _lib.check(_lib.load().dfm_gauss_conv(self.dim, dv, self._h, _p(X), _p(self._Y),
                                              _p(self._W), _p(out), X.shape[0],
                                              self._Y.shape[0], self.precision))
        return out

    def ConvolveGradient(self, X, arg=None):
        """(X) -> (N, Dv, d) list form; (X, alpha) -> (N, d); (X, dim) -> (N, Dv)."""
        self._check_rows()
        X = self._targets(X)
        n, dv, d, ny = X.shape[0], self._W.shape[1], self.dim, self._Y.shape[0]
        L = _lib.load()
        if arg is None:  # CUDAExactKernel.cxx:261-320
            out = np.empty((n, dv, d))
            _lib.check(L.dfm_gauss_gradconv_varlin(d, dv, self._h, _p(X), _p(self._Y), _p(self._W),
                                                   _p(out), n, ny, self.precision))
            return out
        if isinstance(arg, (int, np.integer)):  # CUDAExactKernel.cxx:228-259
            if arg < 0 or arg >= d:
                raise RuntimeError("dimension index out of bounds")
            out = np.empty((n, dv))
            _lib.check(L.dfm_gauss_gradconv(d, dv, self._h, _p(X), _p(self._Y), _p(self._W),
                                            int(arg), _p(out), n, ny, self.precision))
            return out
        alpha = _arr(arg)  # CUDAExactKernel.cxx:165-194
        if alpha.shape != (n, dv):
            raise RuntimeError("alpha must be (N, Dv)")
        out = np.empty((n, d))
        _lib.check(L.dfm_gauss_grad1conv(d, dv, self._h, _p(alpha), _p(X), _p(self._Y),
                                         _p(self._W), _p(out), n, ny, self.precision))
        return out

    def ConvolveAndGradient(self, X, alpha):
        """Convolve(X) and ConvolveGradient(X, alpha) from one pass over the pairs."""
        self._check_rows()
        X = self._targets(X)
        alpha = _arr(alpha)
        n, dv, d, ny = X.shape[0], self._W.shape[1], self.dim, self._Y.shape[0]
        if alpha.shape != (n, dv):
            raise RuntimeError("alpha must be (N, Dv)")
        conv = np.empty((n, dv))
        grad = np.empty((n, d))
        _lib.check(_lib.load().dfm_gauss_conv_grad1(d, dv, self._h, _p(X), _p(alpha), _p(self._Y),
                                                    _p(self._W), _p(conv), _p(grad), n, ny,
                                                    self.precision))
        return conv, grad

    def ConvolveHessian(self, X, *args):
        """(X) -> (N, Dv, d, d), the list form; (X, row, col) -> (N, Dv); (X, k, dp, dq) ->
        (N,) (ExactKernel.cxx:365-468, with its error messages).  One pass over the pairs
        gives every entry; the overloads with indices are slices of it."""
        self._check_rows()
        X = self._targets(X)
        n, dv, d, ny = X.shape[0], self._W.shape[1], self.dim, self._Y.shape[0]
        if len(args) == 2 and (args[0] >= d or args[1] >= d or min(args) < 0):
            raise RuntimeError("Dimension index out of bounds")
        if len(args) == 3:
            if not 0 <= args[0] < dv:
                raise RuntimeError("Invalid weight index")
            if not (0 <= args[1] < d and 0 <= args[2] < d):
                raise RuntimeError("Invalid derivative direction")
        if len(args) not in (0, 2, 3):
            raise TypeError("ConvolveHessian(X), (X, row, col) or (X, k, dp, dq)")
        out = np.empty((n, dv, d, d))
        _lib.check(_lib.load().dfm_gauss_hessianconv(d, dv, self._h, _p(X), _p(self._Y),
                                                     _p(self._W), _p(out), n, ny,
                                                     self.precision))
        if len(args) == 2:
            return np.ascontiguousarray(out[:, :, args[0], args[1]])
        if len(args) == 3:
            return np.ascontiguousarray(out[:, args[0], args[1], args[2]])
        return out

    # CUDAExactKernel.cxx:196-226
    def ConvolveSpecialHessian(self, xi) -> np.ndarray:
        self._check_rows()
        xi = _arr(xi)
        n, d = self._Y.shape
        if xi.shape[0] != n:
            raise RuntimeError("Y and Xi count mismatch")
        dv = self._W.shape[1]
        out = np.empty((n, d))
        _lib.check(_lib.load().dfm_gauss_graddiffconv(d, dv, self._h, _p(self._Y), _p(self._W),
                                                      _p(xi), _p(out), n, self.precision))
        return out
