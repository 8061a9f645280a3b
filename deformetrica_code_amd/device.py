"""Stream-ordered device entry points on torch tensors.

torch provides device memory and the current stream only; every op is one call into
libdfm_hip.so.  Inputs are float64, contiguous, (rows x cols) CUDA tensors on the current
device; results are written into new (or caller-provided) tensors.
"""
from __future__ import annotations

import torch

from . import _lib


def _ptr(t):
    return ctypes_ptr(t) if t is not None else None


def ctypes_ptr(t: torch.Tensor):
    if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
        raise ValueError("expected a contiguous float64 CUDA tensor")
    return t.data_ptr()


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class Workspace:
    """Grow-only device scratch for the device entry points."""

    def __init__(self, device=None):
        self.device = device
        self.buf = None

    def get(self, nbytes: int):
        if nbytes == 0:
            return None, 0
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device or "cuda")
        return self.buf.data_ptr(), self.buf.numel()


_default_ws = {}


def _ws(workspace, device):
    if workspace is not None:
        return workspace
    key = device.index if device.index is not None else torch.cuda.current_device()
    if key not in _default_ws:
        _default_ws[key] = Workspace(device)
    return _default_ws[key]


def conv_grad1(x, alpha, y, w, sigma, precision="fp32c", conv=True, grad=True, stream=None,
               workspace=None):
    """Convolve (N x Dv) and ConvolveGradient(X, alpha) (N x d) in one pass."""
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    cout = torch.empty((n, dv), dtype=torch.float64, device=x.device) if conv is True else conv
    gout = torch.empty((n, d), dtype=torch.float64, device=x.device) if grad is True else grad
    op = _lib.OP_CONV_GRAD1 if (cout is not None and gout is not None) else (
        _lib.OP_CONV if cout is not None else _lib.OP_GRAD1)
    need = L.dfm_workspace_size(op, d, dv, n, y.shape[0], prec)
    wp, wb = _ws(workspace, x.device).get(need)
    _lib.check(L.dfm_conv_grad1(_stream(stream), d, dv, float(sigma), prec, ctypes_ptr(x),
                                _ptr(alpha) if gout is not None else None, n, ctypes_ptr(y),
                                ctypes_ptr(w), y.shape[0], _ptr(cout), _ptr(gout), wp, wb))
    return cout, gout


def convolve(x, y, w, sigma, precision="fp32c", stream=None, workspace=None):
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    out = torch.empty((n, dv), dtype=torch.float64, device=x.device)
    wp, wb = _ws(workspace, x.device).get(L.dfm_workspace_size(_lib.OP_CONV, d, dv, n, y.shape[0], prec))
    _lib.check(L.dfm_conv(_stream(stream), d, dv, float(sigma), prec, ctypes_ptr(x), n,
                          ctypes_ptr(y), ctypes_ptr(w), y.shape[0], ctypes_ptr(out), wp, wb))
    return out


def grad1(x, alpha, y, w, sigma, precision="fp32c", stream=None, workspace=None):
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    out = torch.empty((n, d), dtype=torch.float64, device=x.device)
    wp, wb = _ws(workspace, x.device).get(L.dfm_workspace_size(_lib.OP_GRAD1, d, dv, n, y.shape[0], prec))
    _lib.check(L.dfm_grad1(_stream(stream), d, dv, float(sigma), prec, ctypes_ptr(x),
                           ctypes_ptr(alpha), n, ctypes_ptr(y), ctypes_ptr(w), y.shape[0],
                           ctypes_ptr(out), wp, wb))
    return out


def varlin(x, y, w, sigma, precision="fp32c", stream=None, workspace=None):
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    out = torch.empty((n, dv, d), dtype=torch.float64, device=x.device)
    wp, wb = _ws(workspace, x.device).get(L.dfm_workspace_size(_lib.OP_VARLIN, d, dv, n, y.shape[0], prec))
    _lib.check(L.dfm_varlin(_stream(stream), d, dv, float(sigma), prec, ctypes_ptr(x), n,
                            ctypes_ptr(y), ctypes_ptr(w), y.shape[0], ctypes_ptr(out), wp, wb))
    return out


def graddim(x, y, w, sigma, dim, precision="fp32c", stream=None, workspace=None):
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = w.shape[1]
    out = torch.empty((n, dv), dtype=torch.float64, device=x.device)
    wp, wb = _ws(workspace, x.device).get(L.dfm_workspace_size(_lib.OP_GRADDIM, d, dv, n, y.shape[0], prec))
    _lib.check(L.dfm_graddim(_stream(stream), d, dv, float(sigma), prec, ctypes_ptr(x), n,
                             ctypes_ptr(y), ctypes_ptr(w), y.shape[0], int(dim), ctypes_ptr(out),
                             wp, wb))
    return out


def graddiff(x, beta, eta, sigma, precision="fp32c", stream=None, workspace=None):
    """ConvolveSpecialHessian(eta) with sources x and weights beta (self-interaction)."""
    L = _lib.load()
    prec = _lib.precision_code(precision)
    n, d = x.shape
    dv = beta.shape[1]
    out = torch.empty((n, d), dtype=torch.float64, device=x.device)
    wp, wb = _ws(workspace, x.device).get(L.dfm_workspace_size(_lib.OP_GRADDIFF, d, dv, n, n, prec))
    _lib.check(L.dfm_graddiff(_stream(stream), d, dv, float(sigma), prec, ctypes_ptr(x),
                              ctypes_ptr(beta), ctypes_ptr(eta), n, ctypes_ptr(out), wp, wb))
    return out
