"""Stream-ordered device entry points on device-resident arrays.

``DeviceArray`` is float64 device memory allocated through the library's own HIP runtime
(dfm_malloc); ``Stream`` wraps dfm_stream_create.  Every op is one call into
libdfm_hip.so with device pointers.  No other GPU runtime is involved, so these work in
any process (torch bundles a separate HIP runtime and must not drive the same GPU in the
same process).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class DeviceArray:
    """A C-contiguous array in device memory (float64, or int32 for mesh indices)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._ptr = ctypes.c_void_p()
        _lib.check(_lib.load().dfm_malloc(ctypes.byref(self._ptr), self.nbytes))

    @property
    def ptr(self):
        return self._ptr.value

    @classmethod
    def from_numpy(cls, a, stream=None, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, dtype)
        d.upload(a, stream)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError("shape mismatch %s vs %s" % (a.shape, self.shape))
        _lib.check(_lib.load().dfm_memcpy(self._ptr, a.ctypes.data_as(ctypes.c_void_p),
                                          self.nbytes, 0, _sp(stream)))

    def upload_async(self, a, stream=None):
        """Queue the copy of host array `a` on `stream` without waiting (dfm_memcpy_async):
        `a` must be C-contiguous of this array's dtype and shape, and the caller keeps it
        alive and unchanged until the stream has passed the copy."""
        if not (isinstance(a, np.ndarray) and a.dtype == self.dtype and a.shape == self.shape
                and a.flags.c_contiguous):
            raise ValueError("upload_async needs a C-contiguous %s array of shape %s"
                             % (self.dtype, self.shape))
        _lib.check(_lib.load().dfm_memcpy_async(self._ptr, a.ctypes.data_as(ctypes.c_void_p),
                                                self.nbytes, 0, _sp(stream)))

    def numpy(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _lib.check(_lib.load().dfm_memcpy(out.ctypes.data_as(ctypes.c_void_p), self._ptr,
                                          self.nbytes, 1, _sp(stream)))
        return out

    def free(self):
        if self._ptr.value:
            _lib.load().dfm_free(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self._h = ctypes.c_void_p()
        _lib.check(_lib.load().dfm_stream_create(ctypes.byref(self._h)))
        # the default workspace of calls ordered on this stream (see _ws)
        self.workspace = Workspace()

    @property
    def handle(self):
        return self._h.value

    def synchronize(self):
        _lib.check(_lib.load().dfm_stream_synchronize(self._h))

    def __del__(self):
        try:
            if self._h.value:
                _lib.load().dfm_stream_destroy(self._h)
        except Exception:
            pass


class Workspace:
    """Grow-only device scratch for the device entry points."""

    def __init__(self):
        self.buf = None

    def get(self, nbytes: int):
        if nbytes == 0:
            return None, 0
        if self.buf is None or self.buf.nbytes < nbytes:
            if self.buf is not None:
                self.buf.free()
            self.buf = DeviceArray(((nbytes + 7) // 8,))
        return self.buf.ptr, self.buf.nbytes


def _sp(stream):
    return stream.handle if stream is not None else None


def _p(a):
    return a.ptr if a is not None else None


_default_ws = Workspace()


def _ws(workspace, stream):
    """The workspace of a call: the caller's, else the default one of the call's stream.
    Calls on different streams may run concurrently, so each stream has its own default
    scratch; calls without a stream share the null stream's, which orders them."""
    if workspace is not None:
        return workspace
    return stream.workspace if stream is not None else _default_ws


def device_count() -> int:
    n = ctypes.c_int()
    rc = _lib.load().dfm_device_count(ctypes.byref(n))
    return n.value if rc == _lib.DFM_OK else 0


def conv_grad1(x, alpha, y, w, sigma, precision="fp32c", conv=True, grad=True, stream=None,
               workspace=None):
    """Convolve (N x Dv) and ConvolveGradient(X, alpha) (N x d) in one pass."""
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    cout = DeviceArray((n, dv)) if conv is True else (conv or None)
    gout = DeviceArray((n, d)) if grad is True else (grad or None)
    op = _lib.OP_CONV_GRAD1 if (cout is not None and gout is not None) else (
        _lib.OP_CONV if cout is not None else _lib.OP_GRAD1)
    ws = _ws(workspace, stream)
    wp, wb = ws.get(L.dfm_workspace_size(op, d, dv, n, y.shape[0], prec))
    _lib.check(L.dfm_conv_grad1(_sp(stream), d, dv, float(sigma), prec, x.ptr,
                                _p(alpha) if gout is not None else None, n, y.ptr, w.ptr,
                                y.shape[0], _p(cout), _p(gout), wp, wb))
    return cout, gout


def convolve(x, y, w, sigma, precision="fp32c", stream=None, workspace=None):
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    out = DeviceArray((n, dv))
    wp, wb = _ws(workspace, stream).get(
        L.dfm_workspace_size(_lib.OP_CONV, d, dv, n, y.shape[0], prec))
    _lib.check(L.dfm_conv(_sp(stream), d, dv, float(sigma), prec, x.ptr, n, y.ptr, w.ptr,
                          y.shape[0], out.ptr, wp, wb))
    return out


def grad1(x, alpha, y, w, sigma, precision="fp32c", stream=None, workspace=None):
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    out = DeviceArray((n, d))
    wp, wb = _ws(workspace, stream).get(
        L.dfm_workspace_size(_lib.OP_GRAD1, d, dv, n, y.shape[0], prec))
    _lib.check(L.dfm_grad1(_sp(stream), d, dv, float(sigma), prec, x.ptr, alpha.ptr, n, y.ptr,
                           w.ptr, y.shape[0], out.ptr, wp, wb))
    return out


def varlin(x, y, w, sigma, precision="fp32c", stream=None, workspace=None):
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    out = DeviceArray((n, dv, d))
    wp, wb = _ws(workspace, stream).get(
        L.dfm_workspace_size(_lib.OP_VARLIN, d, dv, n, y.shape[0], prec))
    _lib.check(L.dfm_varlin(_sp(stream), d, dv, float(sigma), prec, x.ptr, n, y.ptr, w.ptr,
                            y.shape[0], out.ptr, wp, wb))
    return out


def graddim(x, y, w, sigma, dim, precision="fp32c", stream=None, workspace=None):
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    out = DeviceArray((n, dv))
    wp, wb = _ws(workspace, stream).get(
        L.dfm_workspace_size(_lib.OP_GRADDIM, d, dv, n, y.shape[0], prec))
    _lib.check(L.dfm_graddim(_sp(stream), d, dv, float(sigma), prec, x.ptr, n, y.ptr, w.ptr,
                             y.shape[0], int(dim), out.ptr, wp, wb))
    return out


def graddiff(x, beta, eta, sigma, precision="fp32c", stream=None, workspace=None):
    """ConvolveSpecialHessian(eta) with sources x and weights beta (self-interaction)."""
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = beta.shape[1]
    out = DeviceArray((n, d))
    wp, wb = _ws(workspace, stream).get(
        L.dfm_workspace_size(_lib.OP_GRADDIFF, d, dv, n, n, prec))
    _lib.check(L.dfm_graddiff(_sp(stream), d, dv, float(sigma), prec, x.ptr, beta.ptr, eta.ptr,
                              n, out.ptr, wp, wb))
    return out


def batch_conv_grad1(x, alpha, y, w, x_offsets, y_offsets, sigma, precision="fp32c", conv=True,
                     grad=True, stream=None, workspace=None):
    """conv_grad1 over a ragged batch: segment s pairs targets x[x_offsets[s]:x_offsets[s+1]]
    with sources y[y_offsets[s]:y_offsets[s+1]] (host int64 offset arrays), one launch."""
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    xo = np.ascontiguousarray(x_offsets, dtype=np.int64)
    yo = np.ascontiguousarray(y_offsets, dtype=np.int64)
    S = len(xo) - 1
    cout = DeviceArray((n, dv)) if conv is True else (conv or None)
    gout = DeviceArray((n, d)) if grad is True else (grad or None)
    op = _lib.OP_CONV_GRAD1 if (cout is not None and gout is not None) else (
        _lib.OP_CONV if cout is not None else _lib.OP_GRAD1)
    ws = _ws(workspace, stream)
    xp, yp = xo.ctypes.data_as(ctypes.c_void_p), yo.ctypes.data_as(ctypes.c_void_p)
    wp, wb = ws.get(L.dfm_batch_workspace(op, d, dv, S, xp, yp, prec))
    _lib.check(L.dfm_batch_conv_grad1(_sp(stream), d, dv, float(sigma), prec, S, xp, yp, x.ptr,
                                      _p(alpha) if gout is not None else None, y.ptr, w.ptr,
                                      _p(cout), _p(gout), wp, wb))
    return cout, gout
