"""Calls on different streams without an explicit workspace: each stream has its own
default scratch, so two calls of different sizes issued back to back on two streams (which
may then run concurrently) both return their own correct sums."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_two_streams_back_to_back(oracle_mod):
    from deformetrica_code_amd import device
    rng = np.random.default_rng(11)
    streams = (device.Stream(), device.Stream())
    cases, dev = [], []
    # the first call (2e10 pairs, milliseconds) is still running when the second is issued
    for (n, m), st in zip(((200000, 100000), (9000, 12000)), streams):
        X, Y = rng.uniform(size=(n, 3)), rng.uniform(size=(m, 3))
        W, A = rng.normal(size=(m, 3)), rng.normal(size=(n, 3))
        cases.append((X, Y, W, A))
        dev.append([device.DeviceArray.from_numpy(v, st) for v in (X, Y, W, A)])
    outs = [device.conv_grad1(dX, dA, dY, dW, 0.05, "fp32c", stream=st)
            for (dX, dY, dW, dA), st in zip(dev, streams)]
    for (X, Y, W, A), (c, g), st in zip(cases, outs, streams):
        st.synchronize()
        rows = np.random.default_rng(3).choice(X.shape[0], 256, replace=False)
        ex = oracle_mod.ExactKernel(3, 0.05)
        ex.SetSources(Y)
        ex.SetWeights(W)
        cref, gref = ex.Convolve(X[rows]), ex.ConvolveGradient(X[rows], A[rows])
        cg, gg = c.numpy(st)[rows], g.numpy(st)[rows]
        assert np.linalg.norm(cg - cref) / np.linalg.norm(cref) < 1e-6
        assert np.linalg.norm(gg - gref) / np.linalg.norm(gref) < 1e-6
