"""Shooting in few launches.  Up to 4,096 control points a single-subject shoot is one
cooperative launch for all its steps (k_shoot_persistent, a grid barrier between steps);
otherwise each step's pair sum over the control points applies the Euler / Heun update of
Diffeos::Shoot (Diffeos.cxx:617-632; the Heun variant of Diffeos.cxx:634-652): on grids of
fewer than 64 target tiles in the same kernel, which also finishes its own source splits
(the last split of each target tile adds up the partials), on larger grids in a launch of
its own after the split reduction.  Checked against the step taken with separate
launches (device conv_grad1, then the update on the host), against the oracle, the two
forms against each other, for single subjects and ragged batches, with and without source
splits, with zero momenta (the reference leaves the points in place,
Diffeos.cxx:609-611), on the matrix-core paths that apply the step in a kernel of its own,
and in explore mode of the autotuner."""
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _shoot(p0, a0, width, T, prec, scheme):
    from deformetrica_code_amd.diffeos import HIPDiffeos
    df = HIPDiffeos(p0.shape[1], prec)
    df.SetKernelWidth(width)
    df.SetNumberOfTimePoints(T)
    df.SetShootingScheme(scheme)
    df.SetStartPositions(p0)
    df.SetStartMomentas(a0)
    df.Update()
    return np.array(df.GetTrajectoryPositions()), np.array(df.GetTrajectoryMomentas())


def _stepwise(p0, a0, width, T, prec, scheme):
    """The same integrator with one device conv_grad1 per field evaluation and the update on
    the host (the separate-launch form of the step)."""
    from deformetrica_code_amd import device
    D = device.DeviceArray.from_numpy

    def field(p, a):
        c, g = device.conv_grad1(D(p), D(a), D(p), D(a), width, prec)
        return c.numpy(), g.numpy()

    dt = 1.0 / (T - 1)
    P, A = [p0], [a0]
    for _ in range(T - 1):
        p, a = P[-1], A[-1]
        dp, da = field(p, a)
        if scheme == "euler":
            P.append(p + dt * dp)
            A.append(a - dt * da)
        else:
            ps, as_ = p + dt * dp, a - dt * da
            dp2, da2 = field(ps, as_)
            P.append(p + 0.5 * dt * dp + 0.5 * dt * dp2)
            A.append(a - 0.5 * dt * da - 0.5 * dt * da2)
    return np.array(P), np.array(A)


@pytest.mark.parametrize("scheme", ["euler", "heun"])
@pytest.mark.parametrize("n", [100, 3000, 4096, 4097, 9000])
def test_fused_step_equals_separate_launches_fp64(scheme, n):
    """n = 100, 3000, 4096: one launch for the whole shoot (1 to 4 LDS chunks of sources);
    4097, 9000: per-step launches with source splits on grids of 65 and more target tiles
    (split reduction and update after the pair sum)."""
    rng = np.random.default_rng(n)
    p0 = rng.uniform(size=(n, 3))
    a0 = rng.normal(scale=0.01, size=(n, 3))
    P, A = _shoot(p0, a0, 0.2, 6, "fp64", scheme)
    Ps, As = _stepwise(p0, a0, 0.2, 6, "fp64", scheme)
    assert rel(P - p0, Ps - p0) < 1e-13
    assert rel(A, As) < 1e-13


@pytest.mark.parametrize("prec,n,tol", [("fp32c", 3000, 1e-6), ("fp32c", 12000, 1e-6),
                                        ("bf16mfma", 12000, 1e-6), ("fp32x", 3000, 1e-6)])
def test_fused_step_fp32_forms_against_oracle(oracle_mod, prec, n, tol):
    """FP32 shoots (12,000 points: Hilbert-sorted targets, and for DFM_BF16MFMA the
    matrix-core kernel with the update in a kernel of its own) against the FP64 oracle."""
    from oracle import deformations as dfo
    rng = np.random.default_rng(n + 1)
    p0 = rng.uniform(size=(n, 3))
    a0 = rng.normal(scale=0.01, size=(n, 3))
    T = 4
    P, A = _shoot(p0, a0, 0.1, T, prec, "euler")
    Po, Ao = dfo.shoot(p0, a0, 0.1, T)
    assert rel(P[-1] - p0, Po[-1] - p0) < tol
    assert rel(A[-1], Ao[-1]) < tol


def test_zero_momenta_gate_in_the_fused_step():
    p0 = np.random.default_rng(3).uniform(size=(3000, 3))
    for scheme in ("euler", "heun"):
        P, A = _shoot(p0, np.zeros_like(p0), 0.2, 5, "fp64", scheme)
        for t in range(5):
            np.testing.assert_array_equal(P[t], p0)
            np.testing.assert_array_equal(A[t], np.zeros_like(p0))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32c", 1e-6)])
def test_one_launch_shoot_equals_per_step_launches(prec, tol):
    """The same 4,096-point shoot as one cooperative launch (dfm_shoot) and as a batch of
    one subject (dfm_batch_shoot: one launch per step)."""
    from deformetrica_code_amd.diffeos import HIPBatchDiffeos
    rng = np.random.default_rng(77)
    p0 = rng.uniform(size=(4096, 3))
    a0 = rng.normal(scale=0.01, size=(4096, 3))
    for scheme in ("euler", "heun"):
        P, A = _shoot(p0, a0, 0.15, 6, prec, scheme)
        b = HIPBatchDiffeos(3, prec)
        b.SetKernelWidth(0.15)
        b.SetNumberOfTimePoints(6)
        b.SetShootingScheme(scheme)
        b.SetStartPositions([p0])
        b.SetStartMomentas([a0])
        b.SetLandmarkPoints([p0[:5]])
        b.Update()
        Pb, Ab = np.array(b.GetTrajectoryPositions()[0]), np.array(b.GetTrajectoryMomentas()[0])
        assert rel(P - p0, Pb - p0) < tol, scheme
        assert rel(A, Ab) < tol, scheme


@pytest.mark.parametrize("sizes", [[700, 1, 2600, 64, 65], [500, 1, 300, 64, 65]])
def test_batched_fused_step_equals_single_subjects(sizes):
    """Ragged batch of subjects (the atlas's shooting): each subject's trajectory equals its
    own single-subject shoot, one of them with zero momenta (gated per segment).  The first
    batch's grid (41 x 5 target tiles) takes the split reduction and the update after the
    pair sum; the second's (8 x 5, split sources) finishes its splits and the step in the
    pair-sum kernel."""
    from deformetrica_code_amd.diffeos import HIPBatchDiffeos
    rng = np.random.default_rng(8)
    p0s = [rng.uniform(size=(s, 3)) for s in sizes]
    a0s = [rng.normal(scale=0.02, size=(s, 3)) for s in sizes]
    a0s[3] = np.zeros((64, 3))
    x0s = [rng.uniform(size=(10, 3)) for _ in sizes]
    for scheme in ("euler", "heun"):
        b = HIPBatchDiffeos(3, "fp64")
        b.SetKernelWidth(0.25)
        b.SetNumberOfTimePoints(5)
        b.SetShootingScheme(scheme)
        b.SetStartPositions(p0s)
        b.SetStartMomentas(a0s)
        b.SetLandmarkPoints(x0s)
        b.Update()
        Pb, Ab = b.GetTrajectoryPositions(), b.GetTrajectoryMomentas()
        for s in range(len(sizes)):
            P, A = _shoot(p0s[s], a0s[s], 0.25, 5, "fp64", scheme)
            assert rel(np.array(Pb[s]) - p0s[s], P - p0s[s]) < 1e-13, (scheme, s)
            assert rel(np.array(Ab[s]), A) < 1e-13, (scheme, s)


def test_fused_step_in_autotuner_explore_mode(oracle_mod):
    """Explore mode runs every candidate plan of the step's pair sum, the matrix-core
    accumulation among them (its step applied by k_step_update): every shoot matches the
    oracle."""
    from deformetrica_code_amd import _lib
    from oracle import deformations as dfo
    rng = np.random.default_rng(21)
    p0 = rng.uniform(size=(5000, 3))   # above the one-launch shoot's 4,096 points
    a0 = rng.normal(scale=0.01, size=(5000, 3))
    Po, Ao = dfo.shoot(p0, a0, 0.1, 3)
    L = _lib.load()
    L.dfm_autotune_reset()
    _lib.autotune(1)
    try:
        for _ in range(25):
            P, A = _shoot(p0, a0, 0.1, 3, "fp32c", "euler")
            assert rel(P[-1] - p0, Po[-1] - p0) < 1e-6
            assert rel(A[-1], Ao[-1]) < 1e-6
        # an explored matrix-core candidate: " R/waves/min_tiles/1:<ms>" with a timing
        assert re.search(r" \d+/\d+/\d+/1:\d", _lib.autotune_report())
    finally:
        _lib.autotune(0)
        L.dfm_autotune_reset()
