"""ORACLE — test infrastructure only.

numpy front-end for ``liboracle.so``, the FP64 C restatement of Deformetrica's CPU
``ExactKernel`` (src/support/kernels/ExactKernel.cxx) and of the generic
``AbstractKernel::ConvolveSpecialHessian`` (src/support/kernels/AbstractKernel.cxx:22-88).
Only the test-suite, the smoke check and the CPU-baseline leg of the benchmark use it, as
the checker.  Matrices are numpy float64 arrays of shape (rows, cols), row-major.

The method names follow the reference's ``AbstractKernel`` API so that parity tests read
like the reference's own ``TestKernelPrecisionCUDA``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile liboracle.so from exact_kernel.c with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i, l, d, p = ctypes.c_int, ctypes.c_long, ctypes.c_double, _D
        L.orc_convolve.argtypes = [i, i, d, p, l, p, p, l, p]
        L.orc_varlin.argtypes = [i, i, d, p, l, p, p, l, p]
        L.orc_grad1.argtypes = [i, i, d, p, p, l, p, p, l, p]
        L.orc_grad_dim.argtypes = [i, i, d, p, l, p, p, l, i, p]
        L.orc_grad_dim.restype = i
        L.orc_hessian_list.argtypes = [i, i, d, p, l, p, p, l, p]
        L.orc_hessian_rc.argtypes = [i, i, d, p, l, p, p, l, i, i, p]
        L.orc_hessian_rc.restype = i
        L.orc_special_hessian.argtypes = [i, d, p, p, p, l, p]
        L.orc_special_hessian.restype = i
        L.orc_special_hessian_graddiff.argtypes = [i, d, p, p, p, l, p]
        L.orc_kernel_matrix.argtypes = [i, d, p, l, p]
        L.orc_convolve_inverse.argtypes = [i, i, d, p, l, p, p]
        L.orc_convolve_inverse.restype = i
        L.orc_convolve_grad1_threaded.argtypes = [i, i, i, d, p, p, l, p, p, l, p, p]
        L.orc_convolve_grad1_threaded.restype = i
        L.orc_mt_stream.argtypes = [ctypes.c_uint32, l, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_reference_fixture.argtypes = [l, i, p]
        _lib = L
    return _lib


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


# Host threads for the row-parallel sums below (each target row's sum is independent of the
# others and computed by the same C loop, so the results do not depend on the split).
_threads = max(1, min(16, os.cpu_count() or 1))


def set_threads(n: int) -> None:
    global _threads
    _threads = max(1, int(n))


def _rows(n, ny, call):
    """call(r0, r1) fills target rows [r0, r1); large sums are split over host threads
    (ctypes releases the GIL during the foreign call)."""
    if _threads == 1 or n < 2 or n * ny < (1 << 22):
        call(0, n)
        return
    parts = min(4 * _threads, n)
    b = np.linspace(0, n, parts + 1).astype(np.int64)
    with ThreadPoolExecutor(min(_threads, parts)) as ex:
        list(ex.map(lambda k: call(int(b[k]), int(b[k + 1])), range(parts)))


def _p(a):
    return a.ctypes.data_as(_D) if a is not None else None


class ExactKernel:
    """FP64 oracle with the reference ``ExactKernel<double, D>`` interface."""

    def __init__(self, dim: int, kernel_width: float = 1.0):
        self.dim = int(dim)
        self.SetKernelWidth(kernel_width)
        self._Y = np.zeros((0, self.dim))
        self._W = np.zeros((0, 1))

    # AbstractKernel.h:85-108
    def SetKernelWidth(self, h: float) -> None:
        self.h = max(float(h), 1e-10)

    def GetKernelWidth(self) -> float:
        return self.h

    def SetSources(self, Y) -> None:
        Y = _c(Y)
        assert Y.ndim == 2 and Y.shape[1] == self.dim
        self._Y = Y.copy()

    def SetWeights(self, W) -> None:
        self._W = _c(W).copy()

    def GetSources(self):
        return self._Y

    def GetWeights(self):
        return self._W

    def _check(self):
        if self._Y.shape[0] != self._W.shape[0]:
            raise RuntimeError("Sources and weights count mismatch")

    def Convolve(self, X):
        self._check()
        X = _c(X)
        dv = self._W.shape[1]
        V = np.empty((X.shape[0], dv))
        Y, W, ny = self._Y, self._W, self._Y.shape[0]
        _rows(X.shape[0], ny, lambda a, b: lib().orc_convolve(
            self.dim, dv, self.h, _p(X[a:b]), b - a, _p(Y), _p(W), ny, _p(V[a:b])))
        return V

    def ConvolveGradient(self, X, arg=None):
        """ConvolveGradient(X) -> list as (N, Dv, d); (X, alpha) -> (N, d); (X, dim) -> (N, Dv)."""
        self._check()
        X = _c(X)
        n, dv, d = X.shape[0], self._W.shape[1], self.dim
        ny = self._Y.shape[0]
        Y, W = self._Y, self._W
        if arg is None:
            G = np.empty((n, dv, d))
            _rows(n, ny, lambda a, b: lib().orc_varlin(
                d, dv, self.h, _p(X[a:b]), b - a, _p(Y), _p(W), ny, _p(G[a:b])))
            return G
        if isinstance(arg, (int, np.integer)):
            if not 0 <= int(arg) < d:
                raise RuntimeError("dimension index out of bounds")
            out = np.empty((n, dv))
            _rows(n, ny, lambda a, b: lib().orc_grad_dim(
                d, dv, self.h, _p(X[a:b]), b - a, _p(Y), _p(W), ny, int(arg), _p(out[a:b])))
            return out
        alpha = _c(arg)
        out = np.empty((n, d))
        _rows(n, ny, lambda a, b: lib().orc_grad1(
            d, dv, self.h, _p(X[a:b]), _p(alpha[a:b]), b - a, _p(Y), _p(W), ny, _p(out[a:b])))
        return out

    def ConvolveHessian(self, X, row=None, col=None):
        self._check()
        X = _c(X)
        n, dv, d = X.shape[0], self._W.shape[1], self.dim
        ny = self._Y.shape[0]
        if row is None:
            H = np.empty((n, dv, d, d))
            lib().orc_hessian_list(d, dv, self.h, _p(X), n, _p(self._Y), _p(self._W), ny, _p(H))
            return H
        out = np.empty((n, dv))
        if lib().orc_hessian_rc(d, dv, self.h, _p(X), n, _p(self._Y), _p(self._W), ny, int(row),
                                int(col), _p(out)) != 0:
            raise RuntimeError("Dimension index out of bounds")
        return out

    def ConvolveSpecialHessian(self, xi):
        """Generic form, AbstractKernel.cxx:22-88."""
        self._check()
        xi = _c(xi)
        n, d = self._Y.shape
        out = np.empty((n, d))
        rc = lib().orc_special_hessian(d, self.h, _p(self._Y), _p(self._W), _p(xi), n, _p(out))
        if rc != 0:
            raise MemoryError("oracle allocation failed")
        return out

    def ConvolveSpecialHessianGradDiff(self, xi):
        """The reference's GPU closed form (CUDAExactKernel.cxx:196-226), on the CPU."""
        self._check()
        xi = _c(xi)
        n, d = self._Y.shape
        out = np.empty((n, d))
        lib().orc_special_hessian_graddiff(d, self.h, _p(self._Y), _p(self._W), _p(xi), n, _p(out))
        return out

    def ComputeKernelMatrix(self, Y):
        Y = _c(Y)
        n = Y.shape[0]
        out = np.empty((n, n))
        lib().orc_kernel_matrix(self.dim, self.h, _p(Y), n, _p(out))
        return out

    def ConvolveInverse(self, X):
        X = _c(X)
        n, dv = X.shape
        V = np.empty((n, dv))
        rc = lib().orc_convolve_inverse(self.dim, dv, self.h, _p(self._Y), n, _p(X), _p(V))
        if rc != 0:
            raise RuntimeError("kernel matrix is singular (rc=%d)" % rc)
        return V


def convolve_grad1_threaded(X, alpha, Y, W, h, nthreads, want_conv=True, want_grad=True):
    """Convolve and ConvolveGradient(X, alpha) on ``nthreads`` host threads."""
    X, alpha, Y, W = _c(X), _c(alpha), _c(Y), _c(W)
    n, d = X.shape
    dv = W.shape[1]
    conv = np.empty((n, dv)) if want_conv else None
    grad = np.empty((n, d)) if want_grad else None
    rc = lib().orc_convolve_grad1_threaded(int(nthreads), d, dv, float(h), _p(X), _p(alpha), n,
                                           _p(Y), _p(W), Y.shape[0], _p(conv), _p(grad))
    if rc != 0:
        raise RuntimeError("thread pool failed (rc=%d)" % rc)
    return conv, grad


def mt19937_stream(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    lib().orc_mt_stream(seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def reference_fixture(rows: int, cols: int) -> np.ndarray:
    """AbstractTestKernelPrecision::generate_random_matrix (.cxx:21-29)."""
    X = np.empty((rows, cols))
    lib().orc_reference_fixture(rows, cols, _p(X))
    return X
