"""MI355X-native Gaussian kernel-sum engine for Deformetrica.

The compute path is libdfm_hip.so (hand-written gfx950 HIP kernels behind the C ABI in
include/dfm_hip.h).  ``HIPExactKernel`` mirrors the reference's exact-kernel plugin
interface; ``device`` exposes the stream-ordered entry points on device arrays that the
library allocates itself.  The callers: ``diffeos`` (shooting, flow, adjoints, parallel
transport), the data terms ``surfaces``, ``curves``, ``volumes`` and ``points``, the models
in ``atlas`` and ``regression``, the ``estimators``, and the file formats in ``io``.
"""
from ._lib import DfmError, load  # noqa: F401
from .kernel import HIPExactKernel  # noqa: F401

__all__ = ["HIPExactKernel", "DfmError", "load"]
