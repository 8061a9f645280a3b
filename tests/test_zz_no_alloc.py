"""dfm_conv_grad1_noalloc (static device scratch, no allocation) against dfm_conv_grad1
(caller workspace), through the C ABI.

Both run the same plan at the same sizes, so their outputs must be bitwise equal.  Two calls
in a row, the second with a smaller input, show that the scratch left over from the first
call does not leak into the second.  A call whose plan needs more than the static scratch
holds is refused with DFM_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run_pair(L, device, x, a, y, w, sigma, prec):
    """(conv, grad) from dfm_conv_grad1 with a workspace and from the no-alloc entry point."""
    from deformetrica_code_amd import _lib
    n, d = x.shape
    m, dv = w.shape
    dx, da, dy, dw = (device.DeviceArray.from_numpy(v) for v in (x, a, y, w))
    c1, g1 = device.DeviceArray((n, dv)), device.DeviceArray((n, d))
    c2, g2 = device.DeviceArray((n, dv)), device.DeviceArray((n, d))
    ws = L.dfm_workspace_size(_lib.OP_CONV_GRAD1, d, dv, n, m, prec)
    work = device.DeviceArray(((ws + 7) // 8,)) if ws else None
    _lib.check(L.dfm_conv_grad1(None, d, dv, sigma, prec, dx.ptr, da.ptr, n, dy.ptr, dw.ptr, m,
                                c1.ptr, g1.ptr, work.ptr if work else None, ws))
    _lib.check(L.dfm_conv_grad1_noalloc(None, d, dv, sigma, prec, dx.ptr, da.ptr, n, dy.ptr,
                                        dw.ptr, m, c2.ptr, g2.ptr))
    return (c1.numpy(), g1.numpy()), (c2.numpy(), g2.numpy())


@pytest.mark.parametrize("prec_name", ["fp64", "fp32c"])
def test_noalloc_equals_workspace_entry_point(prec_name):
    from deformetrica_code_amd import _lib, device
    L = _lib.load()
    prec = _lib.precision_code(prec_name)
    rng = np.random.default_rng(7)
    # first call: 20,000 x 30,000 (Hilbert-sorted FP32 targets, split partials in the
    # scratch); second call smaller: 3,000 x 2,500
    for n, m in ((20000, 30000), (3000, 2500)):
        print("noalloc case", prec_name, n, m, flush=True)
        x = rng.uniform(size=(n, 3))
        a = rng.normal(size=(n, 3))
        y = rng.uniform(size=(m, 3))
        w = rng.normal(size=(m, 3))
        (c1, g1), (c2, g2) = _run_pair(L, device, x, a, y, w, 0.05, prec)
        assert np.array_equal(c1, c2), (prec_name, n, m)
        assert np.array_equal(g1, g2), (prec_name, n, m)
        assert np.linalg.norm(c1) > 0 and np.linalg.norm(g1) > 0


def test_noalloc_refuses_oversized_call():
    from deformetrica_code_amd import _lib, device
    L = _lib.load()
    # 1M targets: the split partials alone exceed the 64 MiB scratch
    n, m = 1 << 20, 4096
    x = device.DeviceArray((n, 3))
    y = device.DeviceArray((m, 3))
    rc = L.dfm_conv_grad1_noalloc(None, 3, 3, 0.05, _lib.FP32C, x.ptr, x.ptr, n, y.ptr, y.ptr, m,
                                  x.ptr, x.ptr)
    assert rc == _lib.DFM_ERR_ARG
    assert b"static scratch" in L.dfm_last_error()
