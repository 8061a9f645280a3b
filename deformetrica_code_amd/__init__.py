"""MI355X-native Gaussian kernel-sum engine for Deformetrica.

The compute path is libdfm_hip.so (hand-written gfx950 HIP kernels behind the C ABI in
include/dfm_hip.h).  ``HIPExactKernel`` mirrors the reference's exact-kernel plugin
interface; ``device`` exposes the stream-ordered entry points on torch tensors.
"""
from ._lib import DfmError, load  # noqa: F401
from .kernel import HIPExactKernel  # noqa: F401

__all__ = ["HIPExactKernel", "DfmError", "load"]
