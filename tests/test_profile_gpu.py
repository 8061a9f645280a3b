"""Profiling of the pair sums (dfm_profile_*): HIP-event kernel time and the library's count
of algorithmic work, single calls and ragged batches."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _read(L):
    tot, cnt = ctypes.c_double(), ctypes.c_long()
    pairs, flops = ctypes.c_double(), ctypes.c_double()
    L.dfm_profile_read(ctypes.byref(tot), ctypes.byref(cnt))
    L.dfm_profile_work(ctypes.byref(pairs), ctypes.byref(flops))
    return tot.value, cnt.value, pairs.value, flops.value


@pytest.mark.parametrize("prec", ["fp64", "fp32c", "bf16mfma"])
def test_profile_counts_pairs_and_flops(prec):
    from deformetrica_code_amd import _lib, device
    L = _lib.load()
    rng = np.random.default_rng(4)
    n, m = 9000, 7000
    D = device.DeviceArray.from_numpy
    X, Y, W, A = (D(rng.uniform(size=(n, 3))), D(rng.uniform(size=(m, 3))),
                  D(rng.normal(size=(m, 3))), D(rng.normal(size=(n, 3))))
    _read(L)
    L.dfm_profile_enable(1)
    try:
        device.conv_grad1(X, A, Y, W, 0.1, prec)
        device.convolve(X, Y, W, 0.1, prec)
        device.batch_conv_grad1(X, A, Y, W, [0, 1000, n], [0, 5000, m], 0.1, prec)
    finally:
        L.dfm_profile_enable(0)
    ms, launches, pairs, flops = _read(L)
    assert launches == 3 and ms > 0
    batch_pairs = 1000 * 5000 + (n - 1000) * (m - 5000)
    assert pairs == 2 * n * m + batch_pairs
    assert flops == 29 * n * m + 16 * n * m + 29 * batch_pairs
    # read and reset: nothing left, and nothing counted while disabled
    device.convolve(X, Y, W, 0.1, prec)
    assert _read(L)[1:] == (0, 0.0, 0.0)
