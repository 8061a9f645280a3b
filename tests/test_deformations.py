"""Pins the oracle's deformation integrators (oracle/deformations.py) on the CPU."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "parallel_transport")


def test_parallel_transport_reference_golden_vectors():
    """TestParallelTransport.cxx:22-83: 2D, exact kernel width 15, 30 time points, the
    regression geodesic shot from CPRegression/MomRegression (explicit Euler, as
    Diffeos::Shoot), MomToTransport transported to t = 1 at target times i/9; the last
    momenta equal GroundTruthMom.txt to 1e-4 absolute."""
    from oracle import deformations as dfo
    cp_reg = dfo.read_dlm(os.path.join(GOLD, "CPRegression.txt"))
    mom_reg = dfo.read_dlm(os.path.join(GOLD, "MomRegression.txt"))
    cp_match = dfo.read_dlm(os.path.join(GOLD, "CPToTransport.txt"))
    mom_match = dfo.read_dlm(os.path.join(GOLD, "MomToTransport.txt"))
    truth = dfo.read_dlm(os.path.join(GOLD, "GroundTruthMom.txt"))
    assert cp_reg.shape == mom_reg.shape == cp_match.shape == mom_match.shape == truth.shape == (144, 2)
    P, A = dfo.shoot(cp_reg, mom_reg, 15.0, 30)
    out = dfo.parallel_transport(P, A, 15.0, 30, mom_match, cp_match, 0.0, [i / 9.0 for i in range(10)])
    assert len(out) == 10
    assert np.max(np.abs(out[-1] - truth)) <= 1e-4


def test_shoot_conserves_hamiltonian_to_first_order():
    from oracle import deformations as dfo
    rng = np.random.default_rng(3)
    p0 = rng.uniform(size=(60, 3))
    a0 = rng.normal(scale=0.05, size=(60, 3))
    errs = []
    for T in (11, 21, 41):
        for scheme in ("euler", "heun"):
            P, A = dfo.shoot(p0, a0, 0.3, T, scheme=scheme)
            h0 = dfo.hamiltonian(P[0], A[0], 0.3)
            errs.append((T, scheme, abs(dfo.hamiltonian(P[-1], A[-1], 0.3) - h0) / h0))
    e = {(T, s): v for T, s, v in errs}
    assert e[(41, "euler")] < e[(11, "euler")] / 3          # Euler: O(dt)
    assert e[(41, "heun")] < e[(11, "heun")] / 10           # Heun: O(dt^2)
    assert e[(41, "heun")] < e[(41, "euler")]


def test_zero_momenta_do_not_move():
    from oracle import deformations as dfo
    p0 = np.random.default_rng(1).uniform(size=(10, 2))
    P, A = dfo.shoot(p0, np.zeros((10, 2)), 0.5, 5)
    assert all(np.array_equal(p, p0) for p in P)
    X = dfo.flow(P, A, p0[:3], 0.5, 5)
    assert all(np.array_equal(x, p0[:3]) for x in X)


def test_reference_adjoint_matches_differences_to_first_order():
    """The reference's continuous adjoint (AdjointEquationsIntegrator) is the gradient of
    E(x_T) = 1/2 |x_T - target|^2 up to O(dt); central differences of the discrete
    shoot+flow objective must agree to that order."""
    from oracle import deformations as dfo
    rng = np.random.default_rng(7)
    n, m, width = 8, 12, 0.6
    p0 = rng.uniform(size=(n, 2))
    a0 = rng.normal(scale=0.2, size=(n, 2))
    x0 = rng.uniform(size=(m, 2))
    target = x0 + rng.normal(scale=0.1, size=(m, 2))

    def objective(p, a, T):
        P, A = dfo.shoot(p, a, width, T)
        X = dfo.flow(P, A, x0, width, T)
        return 0.5 * np.sum((X[-1] - target) ** 2), (P, A, X)

    rel = []
    for T in (11, 41):
        E, (P, A, X) = objective(p0, a0, T)
        gp, ga, _ = dfo.adjoint_reference(P, A, X, X[-1] - target, width, T)
        eps = 1e-6
        fd_a = np.zeros_like(a0)
        for i in range(n):
            for k in range(2):
                ap, am = a0.copy(), a0.copy()
                ap[i, k] += eps
                am[i, k] -= eps
                fd_a[i, k] = (objective(p0, ap, T)[0] - objective(p0, am, T)[0]) / (2 * eps)
        rel.append(np.linalg.norm(ga - fd_a) / np.linalg.norm(fd_a))
    assert rel[1] < rel[0] / 2.5     # first-order agreement
    assert rel[1] < 0.05
