"""Ragged batches of subjects: one launch per kernel sum for all subjects, against the
single-subject entry points (themselves pinned against the oracle), per-subject
zero-momenta exits, and HIP-graph capture and replay of batched shooting and flow."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _offsets(counts):
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    off[1:] = np.cumsum(counts)
    return off


@pytest.mark.parametrize("prec", ["fp64", "fp32c", "fp32x", "bf16mfma"])
def test_batch_conv_grad1_matches_per_segment_calls(prec):
    from deformetrica_code_amd import device
    rng = np.random.default_rng(1)
    nx = [300, 0, 1000, 77, 5000, 1]
    ny = [400, 200, 0, 1500, 9000, 9]     # DFM_BF16MFMA batches need 2,048 sources somewhere
    xo, yo = _offsets(nx), _offsets(ny)
    X = rng.uniform(size=(xo[-1], 3))
    A = rng.normal(size=(xo[-1], 3))
    Y = rng.uniform(size=(yo[-1], 3))
    W = rng.normal(size=(yo[-1], 3))
    sigma = 0.1
    dX, dA, dY, dW = (device.DeviceArray.from_numpy(v) for v in (X, A, Y, W))
    conv, grad = device.batch_conv_grad1(dX, dA, dY, dW, xo, yo, sigma, prec)
    conv, grad = conv.numpy(), grad.numpy()
    for s in range(len(nx)):
        xs, ys = slice(xo[s], xo[s + 1]), slice(yo[s], yo[s + 1])
        if nx[s] == 0:
            continue
        if ny[s] == 0:
            np.testing.assert_array_equal(conv[xs], 0.0)
            np.testing.assert_array_equal(grad[xs], 0.0)
            continue
        c1, g1 = device.conv_grad1(device.DeviceArray.from_numpy(X[xs]),
                                   device.DeviceArray.from_numpy(A[xs]),
                                   device.DeviceArray.from_numpy(Y[ys]),
                                   device.DeviceArray.from_numpy(W[ys]), sigma, "fp64")
        tol = 1e-12 if prec == "fp64" else 1e-6
        assert rel(conv[xs], c1.numpy()) < tol, s
        assert rel(grad[xs], g1.numpy()) < tol, s


def _zorder(x, bits=10):
    """Z-order keys of points in [0, 1]^3 (bit-interleaved cell indices)."""
    q = np.clip((x * (1 << bits)).astype(np.int64), 0, (1 << bits) - 1)
    key = np.zeros(len(x), dtype=np.int64)
    for b in range(bits):
        for k in range(3):
            key |= ((q[:, k] >> b) & 1) << (3 * b + k)
    return key


@pytest.mark.parametrize("op", ["conv", "grad1", "conv_grad1"])
def test_bf16_batch_on_compact_segments_matches_fp64(op):
    """DFM_BF16MFMA batches in the caller's order: segments whose targets are in Z (Morton)
    order (most groups of 256 compact: the matrix-core path) and one in random order (every
    group wide: the FP32C code inside the same kernel), Dv = 3 and 6, against the FP64 batch
    to 1e-6 norm-wise per segment."""
    from deformetrica_code_amd import device
    rng = np.random.default_rng(5)
    nx = [9000, 300, 0, 2500, 257]
    ny = [9000, 4000, 100, 0, 3000]
    xo, yo = _offsets(nx), _offsets(ny)
    X = rng.uniform(size=(xo[-1], 3))
    for s in (0, 3, 4):
        seg = X[xo[s]:xo[s + 1]]
        X[xo[s]:xo[s + 1]] = seg[np.argsort(_zorder(seg), kind="stable")]
    Y = rng.uniform(size=(yo[-1], 3))
    sigma = 0.1
    for dv in (3, 6):
        W = rng.normal(size=(yo[-1], dv))
        A = rng.normal(size=(xo[-1], dv))
        dX, dA, dY, dW = (device.DeviceArray.from_numpy(v) for v in (X, A, Y, W))
        out = {}
        for prec in ("fp64", "bf16mfma"):
            c, g = device.batch_conv_grad1(dX, dA, dY, dW, xo, yo, sigma, prec,
                                           conv=op != "grad1", grad=op != "conv")
            out[prec] = tuple(v.numpy() for v in (c, g) if v is not None)
        for s in range(len(nx)):
            xs = slice(xo[s], xo[s + 1])
            if nx[s] == 0:
                continue
            for a, b in zip(out["bf16mfma"], out["fp64"]):
                if ny[s] == 0:
                    np.testing.assert_array_equal(a[xs], 0.0)
                else:
                    assert rel(a[xs], b[xs]) < 1e-6, (op, dv, s, rel(a[xs], b[xs]))


@pytest.mark.parametrize("nx,ny", [([9000, 300, 2500], [2047, 1000, 512]),
                                   ([4095, 300, 2048], [5000, 3000, 2048])])
def test_bf16_batch_below_the_size_bounds_equals_fp32c(nx, ny):
    """A DFM_BF16MFMA batch whose largest segment has fewer than 2,048 sources (the
    matrix-core kernel's per-group set-up does not pay off over few source tiles) or fewer
    than 4,096 targets runs the DFM_FP32C kernels: equal bit for bit."""
    from deformetrica_code_amd import device
    rng = np.random.default_rng(8)
    xo, yo = _offsets(nx), _offsets(ny)
    X, A = rng.uniform(size=(xo[-1], 3)), rng.normal(size=(xo[-1], 3))
    Y, W = rng.uniform(size=(yo[-1], 3)), rng.normal(size=(yo[-1], 3))
    dX, dA, dY, dW = (device.DeviceArray.from_numpy(v) for v in (X, A, Y, W))
    out = {}
    for prec in ("fp32c", "bf16mfma"):
        c, g = device.batch_conv_grad1(dX, dA, dY, dW, xo, yo, 0.1, prec)
        out[prec] = (c.numpy(), g.numpy())
    for a, b in zip(out["bf16mfma"], out["fp32c"]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("prec", ["fp32c", "fp32x"])
@pytest.mark.parametrize("m", [1, 100, 4096, 50000])
def test_small_grid_form_matches_fp64(prec, m):
    """FP32 calls on 32-127 target workgroups run sixteen wavefronts per workgroup on the
    same 64 targets, their source splits added in LDS: single calls of 4,096 targets and a
    batch of two segments (2,500 and 4,000 targets) against FP64, Convolve, Grad1 and the
    fused sum, norm-wise 1e-6."""
    from deformetrica_code_amd import device
    rng = np.random.default_rng(m)
    n = 4096
    X, A = rng.uniform(size=(n, 3)), rng.normal(size=(n, 3))
    Y, W = rng.uniform(size=(m, 3)), rng.normal(size=(m, 3))
    dX, dA, dY, dW = (device.DeviceArray.from_numpy(v) for v in (X, A, Y, W))
    for conv, grad in ((True, True), (True, False), (False, True)):
        ref = device.conv_grad1(dX, dA, dY, dW, 0.1, "fp64", conv=conv, grad=grad)
        got = device.conv_grad1(dX, dA, dY, dW, 0.1, prec, conv=conv, grad=grad)
        for a, b in zip(got, ref):
            if b is not None:
                assert rel(a.numpy(), b.numpy()) < 1e-6, (conv, grad)
    xo, yo = np.array([0, 2500, n]), np.array([0, m // 3, m])
    ref = device.batch_conv_grad1(dX, dA, dY, dW, xo, yo, 0.1, "fp64")
    got = device.batch_conv_grad1(dX, dA, dY, dW, xo, yo, 0.1, prec)
    for s in range(2):
        xs = slice(xo[s], xo[s + 1])
        for a, b in zip(got, ref):
            a, b = a.numpy()[xs], b.numpy()[xs]
            if yo[s + 1] == yo[s]:
                np.testing.assert_array_equal(a, 0.0)
            else:
                assert rel(a, b) < 1e-6, s


def _subjects(seed, counts_cp, counts_lm, d=3, scale=0.2):
    rng = np.random.default_rng(seed)
    p0 = [rng.uniform(size=(n, d)) for n in counts_cp]
    a0 = [rng.normal(scale=scale, size=(n, d)) for n in counts_cp]
    x0 = [rng.uniform(size=(m, d)) for m in counts_lm]
    return p0, a0, x0


def _single(d, width, T, p0, a0, x0, prec, scheme, improved):
    from deformetrica_code_amd.diffeos import HIPDiffeos
    df = HIPDiffeos(d, prec)
    df.SetKernelWidth(width)
    df.SetNumberOfTimePoints(T)
    df.SetShootingScheme(scheme)
    if not improved:
        df.UseStandardEuler()
    df.SetStartPositions(p0)
    df.SetStartMomentas(a0)
    df.SetLandmarkPoints(x0)
    return df


def _batch(d, width, T, p0, a0, x0, prec, scheme, improved):
    from deformetrica_code_amd.diffeos import HIPBatchDiffeos
    df = HIPBatchDiffeos(d, prec)
    df.SetKernelWidth(width)
    df.SetNumberOfTimePoints(T)
    df.SetShootingScheme(scheme)
    if not improved:
        df.UseStandardEuler()
    df.SetStartPositions(p0)
    df.SetStartMomentas(a0)
    df.SetLandmarkPoints(x0)
    return df


@pytest.mark.parametrize("scheme,improved", [("euler", True), ("heun", False)])
def test_batch_shoot_flow_and_adjoint_match_single_subject_calls(scheme, improved):
    d, width, T = 3, 0.3, 9
    cps, lms = [40, 130, 7, 64], [200, 55, 300, 1]
    p0, a0, x0 = _subjects(2, cps, lms)
    B = _batch(d, width, T, p0, a0, x0, "fp64", scheme, improved)
    B.Update()
    rng = np.random.default_rng(3)
    thetas = [rng.normal(size=x.shape) for x in x0]
    gps, gas, gxs = B.GradientOfTerminalCost(thetas, checkpoint_every=3)
    Ps, As, Xs = B.GetTrajectoryPositions(), B.GetTrajectoryMomentas(), B.GetLandmarkPointsTrajectory()
    for s in range(len(cps)):
        S = _single(d, width, T, p0[s], a0[s], x0[s], "fp64", scheme, improved)
        S.Update()
        assert rel(Ps[s], np.array(S.GetTrajectoryPositions())) < 1e-13
        assert rel(As[s], np.array(S.GetTrajectoryMomentas())) < 1e-12
        assert rel(Xs[s], np.array(S.GetLandmarkPointsTrajectory())) < 1e-13
        gp, ga, gx = S.GradientOfTerminalCost(thetas[s], checkpoint_every=3)
        assert rel(gps[s], gp) < 1e-11 and rel(gas[s], ga) < 1e-11 and rel(gxs[s], gx) < 1e-11


def test_batch_fp32c_within_1e6_of_fp64():
    d, width, T = 3, 0.2, 11
    p0, a0, x0 = _subjects(4, [500, 900], [1000, 300], scale=0.01)
    out = {}
    for prec in ("fp64", "fp32c"):
        B = _batch(d, width, T, p0, a0, x0, prec, "euler", True)
        B.Update()
        out[prec] = (B.GetTrajectoryPositions(), B.GetLandmarkPointsTrajectory())
    for s in range(2):
        assert rel(out["fp32c"][0][s][-1] - p0[s], out["fp64"][0][s][-1] - p0[s]) < 1e-6
        assert rel(out["fp32c"][1][s][-1] - x0[s], out["fp64"][1][s][-1] - x0[s]) < 1e-6


def test_zero_momenta_exit_is_per_subject():
    p0, a0, x0 = _subjects(5, [30, 30, 30], [40, 40, 40])
    a0[1] = np.full_like(a0[1], 1e-22)
    B = _batch(3, 0.4, 6, p0, a0, x0, "fp64", "euler", True)
    B.Update()
    P, X = B.GetTrajectoryPositions(), B.GetLandmarkPointsTrajectory()
    for t in range(6):
        np.testing.assert_array_equal(P[1][t], p0[1])
        np.testing.assert_array_equal(X[1][t], x0[1])
    assert np.abs(P[0][-1] - p0[0]).max() > 1e-3
    assert np.abs(X[2][-1] - x0[2]).max() > 1e-3


def test_graph_capture_replays_batched_shoot_and_flow():
    """Capture dfm_batch_shoot + dfm_batch_flow once, replay it after changing the momenta
    in place, with eager work in between; every replay equals the eager result bitwise."""
    from deformetrica_code_amd import _lib
    from deformetrica_code_amd.device import DeviceArray, Stream
    L = _lib.load()
    d, T, width, prec = 3, 7, 0.3, _lib.FP32C
    cps, lms = [64, 200, 33], [500, 120, 260]
    p0, a0, x0 = _subjects(6, cps, lms)
    cp, lm = _offsets(cps), _offsets(lms)
    cpp, lmp = cp.ctypes.data_as(ctypes.c_void_p), lm.ctypes.data_as(ctypes.c_void_p)
    P0, A0, X0 = (np.concatenate(v) for v in (p0, a0, x0))
    n, m, S = P0.shape[0], X0.shape[0], len(cps)
    st = Stream()
    dp0, da0, dx0 = (DeviceArray.from_numpy(v, st) for v in (P0, A0, X0))
    P, A, X = DeviceArray((T, n, d)), DeviceArray((T, n, d)), DeviceArray((T, m, d))
    w1 = DeviceArray(((L.dfm_batch_shoot_workspace(d, S, cpp, prec) + 7) // 8,))
    w2 = DeviceArray(((L.dfm_batch_flow_workspace(d, S, cpp, lmp, prec) + 7) // 8,))

    def issue():
        _lib.check(L.dfm_batch_shoot(st.handle, d, width, prec, _lib.EULER, S, cpp, dp0.ptr, da0.ptr,
                                     T, 0.0, 1.0, P.ptr, A.ptr, w1.ptr, w1.nbytes))
        _lib.check(L.dfm_batch_flow(st.handle, d, width, prec, 1, S, cpp, lmp, P.ptr, A.ptr, dx0.ptr,
                                    T, 0.0, 1.0, X.ptr, w2.ptr, w2.nbytes))

    def eager():
        issue()
        st.synchronize()
        return P.numpy(st), X.numpy(st)

    ref = eager()
    _lib.check(L.dfm_graph_begin_capture(st.handle))
    issue()
    g = ctypes.c_void_p()
    _lib.check(L.dfm_graph_end_capture(st.handle, ctypes.byref(g)))
    try:
        rng = np.random.default_rng(7)
        for r in range(4):
            if r > 0:
                da0.upload(A0 * (1.0 + 0.1 * r) + rng.normal(scale=1e-3, size=A0.shape), st)
                ref = eager()
                P.upload(np.zeros((T, n, d)), st)
                X.upload(np.zeros((T, m, d)), st)
            _lib.check(L.dfm_graph_launch(g, st.handle))
            st.synchronize()
            np.testing.assert_array_equal(P.numpy(st), ref[0])
            np.testing.assert_array_equal(X.numpy(st), ref[1])
    finally:
        _lib.check(L.dfm_graph_destroy(g))
