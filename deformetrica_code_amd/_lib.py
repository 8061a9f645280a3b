"""ctypes binding of libdfm_hip.so (the C ABI declared in include/dfm_hip.h).

The library is the only compute path: if it is missing this module raises instead of
falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdfm_hip.so")

DFM_OK, DFM_ERR_ARG, DFM_ERR_MISMATCH, DFM_ERR_DIM, DFM_ERR_INDEX, DFM_ERR_HIP, DFM_ERR_NOMEM = range(7)
FP64, FP32C, FP32X, BF16MFMA = 0, 1, 2, 3
(OP_CONV, OP_GRAD1, OP_CONV_GRAD1, OP_VARLIN, OP_GRADDIM, OP_GRADDIFF, OP_SHOOT_VJP, OP_HESSIAN,
 OP_VARIFOLD, OP_VARIFOLD_CONV) = range(10)
EULER, HEUN = 0, 1
CURRENT, VARIFOLD = 0, 1
ADJOINT_DISCRETE, ADJOINT_CONTINUOUS = 0, 1
ADJOINTS = {"discrete": ADJOINT_DISCRETE, "continuous": ADJOINT_CONTINUOUS}

PRECISIONS = {"fp64": FP64, "fp32c": FP32C, "fp32x": FP32X, "bf16mfma": BF16MFMA}

# DFM_ABI_VERSION of include/dfm_hip.h, the header these bindings follow
ABI_VERSION = 4


class DfmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__("dfm error %d: %s" % (code, message))
        self.code = code


_lib = None

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)


def load():
    """Load libdfm_hip.so, raising if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libdfm_hip.so is not built (%s); run `python -m deformetrica_code_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    i, l, d, sz = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_size_t
    L.dfm_abi_version.restype = i
    L.dfm_last_error.restype = ctypes.c_char_p
    if L.dfm_abi_version() != ABI_VERSION:
        raise RuntimeError("libdfm_hip.so has ABI version %d, these bindings expect %d; rebuild it"
                           % (L.dfm_abi_version(), ABI_VERSION))
    L.dfm_workspace_size.argtypes = [i, i, i, l, l, i]
    L.dfm_workspace_size.restype = sz
    vp = ctypes.c_void_p
    L.dfm_device_count.argtypes = [ctypes.POINTER(i)]
    L.dfm_malloc.argtypes = [ctypes.POINTER(vp), sz]
    L.dfm_free.argtypes = [vp]
    L.dfm_memcpy.argtypes = [vp, vp, sz, i, vp]
    L.dfm_memcpy_async.argtypes = [vp, vp, sz, i, vp]
    L.dfm_stream_create.argtypes = [ctypes.POINTER(vp)]
    L.dfm_stream_destroy.argtypes = [vp]
    L.dfm_stream_synchronize.argtypes = [vp]
    for name in ("dfm_device_count", "dfm_malloc", "dfm_free", "dfm_memcpy", "dfm_memcpy_async",
                 "dfm_stream_create",
                 "dfm_stream_destroy", "dfm_stream_synchronize"):
        getattr(L, name).restype = i
    # integrators
    L.dfm_shoot_workspace.argtypes = [i, l, i]
    L.dfm_shoot_workspace.restype = sz
    L.dfm_shoot.argtypes = [vp, i, d, i, i, vp, vp, l, i, d, d, vp, vp, vp, sz]
    L.dfm_flow_workspace.argtypes = [i, l, l, i]
    L.dfm_flow_workspace.restype = sz
    L.dfm_flow.argtypes = [vp, i, d, i, i, vp, vp, l, vp, l, i, d, d, vp, vp, sz]
    L.dfm_shoot_vjp.argtypes = [vp, i, d, i, vp, vp, vp, vp, l, vp, vp, vp, sz]
    L.dfm_adjoint_workspace.argtypes = [i, l, l, i, i, i]
    L.dfm_adjoint_workspace.restype = sz
    L.dfm_shoot_flow_adjoint.argtypes = [vp, i, d, i, i, i, vp, vp, l, vp, l, i, d, d, i, vp, vp,
                                         vp, vp, vp, sz]
    L.dfm_shoot_flow_adjoint_equations.argtypes = L.dfm_shoot_flow_adjoint.argtypes
    L.dfm_shoot_flow_adjoint_jumps.argtypes = [vp, i, d, i, i, i, vp, vp, l, vp, l, i, d, d, i,
                                               i, vp, vp, vp, vp, vp, vp, sz]
    for name in ("dfm_shoot", "dfm_flow", "dfm_shoot_vjp", "dfm_shoot_flow_adjoint",
                 "dfm_shoot_flow_adjoint_equations", "dfm_shoot_flow_adjoint_jumps"):
        getattr(L, name).restype = i
    # ragged batches and graphs
    L.dfm_batch_workspace.argtypes = [i, i, i, i, vp, vp, i]
    L.dfm_batch_workspace.restype = sz
    L.dfm_batch_conv_grad1.argtypes = [vp, i, i, d, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz]
    L.dfm_batch_shoot_workspace.argtypes = [i, i, vp, i]
    L.dfm_batch_shoot_workspace.restype = sz
    L.dfm_batch_shoot.argtypes = [vp, i, d, i, i, i, vp, vp, vp, i, d, d, vp, vp, vp, sz]
    L.dfm_batch_flow_workspace.argtypes = [i, i, vp, vp, i]
    L.dfm_batch_flow_workspace.restype = sz
    L.dfm_batch_flow.argtypes = [vp, i, d, i, i, i, vp, vp, vp, vp, vp, i, d, d, vp, vp, sz]
    L.dfm_batch_adjoint_workspace.argtypes = [i, i, vp, vp, i, i, i]
    L.dfm_batch_adjoint_workspace.restype = sz
    L.dfm_batch_shoot_flow_adjoint.argtypes = [vp, i, d, i, i, i, i, vp, vp, vp, vp, vp, i, d, d,
                                               i, vp, vp, vp, vp, vp, sz]
    L.dfm_graph_begin_capture.argtypes = [vp]
    L.dfm_graph_end_capture.argtypes = [vp, ctypes.POINTER(vp)]
    L.dfm_graph_launch.argtypes = [vp, vp]
    L.dfm_graph_destroy.argtypes = [vp]
    for name in ("dfm_batch_conv_grad1", "dfm_batch_shoot", "dfm_batch_flow",
                 "dfm_batch_shoot_flow_adjoint", "dfm_graph_begin_capture",
                 "dfm_graph_end_capture", "dfm_graph_launch", "dfm_graph_destroy"):
        getattr(L, name).restype = i
    # kernel matrix, ConvolveInverse, parallel transport
    L.dfm_kernel_matrix.argtypes = [vp, i, d, vp, l, vp]
    L.dfm_convolve_inverse_workspace.argtypes = [i, i, l]
    L.dfm_convolve_inverse_workspace.restype = sz
    L.dfm_convolve_inverse.argtypes = [vp, i, i, d, vp, l, vp, vp, vp, sz]
    L.dfm_parallel_transport_workspace.argtypes = [i, l, l, i]
    L.dfm_parallel_transport_workspace.restype = sz
    L.dfm_parallel_transport.argtypes = [vp, i, d, i, vp, vp, l, i, d, d, vp, vp, l, d, vp, i, vp,
                                         vp, sz]
    for name in ("dfm_kernel_matrix", "dfm_convolve_inverse", "dfm_parallel_transport"):
        getattr(L, name).restype = i
    # deterministic atlas
    L.dfm_atlas_create.argtypes = [vp, ctypes.POINTER(vp)]
    L.dfm_atlas_destroy.argtypes = [vp]
    L.dfm_atlas_device_bytes.argtypes = [vp]
    L.dfm_atlas_device_bytes.restype = sz
    L.dfm_atlas_gradient.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    for name in ("dfm_atlas_create", "dfm_atlas_destroy", "dfm_atlas_gradient"):
        getattr(L, name).restype = i
    # surface data terms
    L.dfm_mesh_incidence.argtypes = [vp, l, l, vp, vp]
    L.dfm_surface_workspace.argtypes = [i, l, l, i, i]
    L.dfm_surface_workspace.restype = sz
    L.dfm_surface_match.argtypes = [vp, i, d, i, vp, l, vp, l, vp, l, vp, l, vp, vp, sz]
    L.dfm_surface_match_gradient.argtypes = [vp, i, d, i, vp, l, vp, l, vp, vp, vp, l, vp, l, vp,
                                             vp, sz]
    L.dfm_samples_workspace.argtypes = [i, l, l, i, i]
    L.dfm_samples_workspace.restype = sz
    L.dfm_samples_match_gradient.argtypes = [vp, i, d, i, vp, vp, l, vp, vp, l, vp, vp, vp, vp, sz]
    # polyline data terms
    L.dfm_cell_incidence.argtypes = [vp, l, i, l, vp, vp]
    L.dfm_curve_workspace.argtypes = [i, i, l, l, i, i]
    L.dfm_curve_workspace.restype = sz
    L.dfm_curve_match.argtypes = [vp, i, i, d, i, vp, l, vp, l, vp, l, vp, l, vp, vp, sz]
    L.dfm_curve_match_gradient.argtypes = [vp, i, i, d, i, vp, l, vp, l, vp, vp, vp, l, vp, l,
                                           vp, vp, sz]
    # volume-mesh data terms
    L.dfm_volume_workspace.argtypes = [i, l, l, i, i]
    L.dfm_volume_workspace.restype = sz
    L.dfm_volume_match.argtypes = [vp, i, d, i, vp, l, vp, l, vp, l, vp, l, vp, vp, sz]
    L.dfm_volume_match_gradient.argtypes = [vp, i, d, i, vp, l, vp, l, vp, vp, vp, l, vp, l, vp,
                                            vp, sz]
    # point-cloud and landmark data terms
    L.dfm_pointcloud_workspace.argtypes = [i, l, l, i, i]
    L.dfm_pointcloud_workspace.restype = sz
    L.dfm_pointcloud_match.argtypes = [vp, i, d, i, vp, vp, l, vp, vp, l, vp, vp, sz]
    L.dfm_pointcloud_match_gradient.argtypes = [vp, i, d, i, vp, vp, l, vp, vp, l, vp, vp, sz]
    L.dfm_landmark_workspace.argtypes = [i, l]
    L.dfm_landmark_workspace.restype = sz
    L.dfm_landmark_match.argtypes = [vp, i, vp, vp, l, vp, vp, sz]
    L.dfm_landmark_match_gradient.argtypes = [vp, i, vp, vp, l, vp]
    for name in ("dfm_mesh_incidence", "dfm_surface_match", "dfm_surface_match_gradient",
                 "dfm_samples_match_gradient", "dfm_cell_incidence", "dfm_curve_match",
                 "dfm_curve_match_gradient", "dfm_volume_match", "dfm_volume_match_gradient",
                 "dfm_pointcloud_match", "dfm_pointcloud_match_gradient", "dfm_landmark_match",
                 "dfm_landmark_match_gradient"):
        getattr(L, name).restype = i
    L.dfm_autotune.argtypes = [i]
    L.dfm_autotune.restype = i
    L.dfm_autotune_reset.argtypes = []
    L.dfm_autotune_reset.restype = i
    L.dfm_autotune_report.argtypes = [ctypes.c_char_p, sz]
    L.dfm_autotune_report.restype = i
    L.dfm_profile_enable.argtypes = [i]
    L.dfm_profile_enable.restype = i
    L.dfm_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
    L.dfm_profile_read.restype = i
    L.dfm_profile_work.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.dfm_profile_work.restype = i
    L.dfm_flops_per_pair.argtypes = [i, i, i]
    L.dfm_flops_per_pair.restype = ctypes.c_double
    # device entry points
    L.dfm_conv.argtypes = [_P, i, i, d, i, _P, l, _P, _P, l, _P, _P, sz]
    L.dfm_grad1.argtypes = [_P, i, i, d, i, _P, _P, l, _P, _P, l, _P, _P, sz]
    L.dfm_conv_grad1.argtypes = [_P, i, i, d, i, _P, _P, l, _P, _P, l, _P, _P, _P, sz]
    L.dfm_conv_grad1_noalloc.argtypes = [_P, i, i, d, i, _P, _P, l, _P, _P, l, _P, _P]
    L.dfm_varlin.argtypes = [_P, i, i, d, i, _P, l, _P, _P, l, _P, _P, sz]
    L.dfm_graddim.argtypes = [_P, i, i, d, i, _P, l, _P, _P, l, i, _P, _P, sz]
    L.dfm_graddiff.argtypes = [_P, i, i, d, i, _P, _P, _P, l, _P, _P, sz]
    L.dfm_hessian.argtypes = [_P, i, i, d, i, _P, l, _P, _P, l, _P, _P, sz]
    # host-pointer wrappers
    L.dfm_gauss_conv.argtypes = [i, i, d, _D, _D, _D, _D, i, i, i]
    L.dfm_gauss_grad1conv.argtypes = [i, i, d, _D, _D, _D, _D, _D, i, i, i]
    L.dfm_gauss_gradconv.argtypes = [i, i, d, _D, _D, _D, ctypes.c_uint, _D, i, i, i]
    L.dfm_gauss_gradconv_varlin.argtypes = [i, i, d, _D, _D, _D, _D, i, i, i]
    L.dfm_gauss_graddiffconv.argtypes = [i, i, d, _D, _D, _D, _D, i, i]
    L.dfm_gauss_conv_grad1.argtypes = [i, i, d, _D, _D, _D, _D, _D, _D, i, i, i]
    L.dfm_gauss_hessianconv.argtypes = [i, i, d, _D, _D, _D, _D, i, i, i]
    for name in ("dfm_conv", "dfm_grad1", "dfm_conv_grad1", "dfm_conv_grad1_noalloc",
                 "dfm_varlin", "dfm_graddim",
                 "dfm_graddiff", "dfm_hessian", "dfm_gauss_conv", "dfm_gauss_grad1conv",
                 "dfm_gauss_gradconv", "dfm_gauss_gradconv_varlin", "dfm_gauss_graddiffconv",
                 "dfm_gauss_conv_grad1", "dfm_gauss_hessianconv"):
        getattr(L, name).restype = i
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != DFM_OK:
        msg = load().dfm_last_error() or b""
        raise DfmError(rc, msg.decode(errors="replace"))


def adjoint_code(adjoint) -> int:
    if isinstance(adjoint, int):
        return adjoint
    try:
        return ADJOINTS[adjoint]
    except KeyError:
        raise ValueError(f"unknown adjoint form {adjoint!r} (discrete or continuous)") from None


def precision_code(precision) -> int:
    if isinstance(precision, int):
        return precision
    try:
        return PRECISIONS[precision.lower()]
    except KeyError:
        raise ValueError("precision must be one of %s" % sorted(PRECISIONS)) from None


def autotune(mode: int) -> None:
    """0 heuristics, 1 explore then keep the fastest plan per shape, 2 tuned plans only."""
    check(load().dfm_autotune(int(mode)))


def autotune_report() -> str:
    L = load()
    n = L.dfm_autotune_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.dfm_autotune_report(buf, n + 1)
    return buf.value.decode()
