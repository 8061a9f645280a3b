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


def test_flow_landmark_on_a_lone_control_point_follows_it_exactly():
    """One control point p with momentum a moves at K(p, p) a = a with constant momentum;
    a landmark placed on it sees velocity a at every stage of either scheme, so it stays on
    the control point exactly."""
    from oracle import deformations as dfo
    p0 = np.array([[0.3, -0.2, 0.5]])
    a0 = np.array([[0.4, 0.1, -0.7]])
    for scheme in ("euler", "heun"):
        P, A = dfo.shoot(p0, a0, 0.7, 9, scheme=scheme)
        for improved in (True, False):
            X = dfo.flow(P, A, p0, 0.7, 9, improved_euler=improved)
            for t in range(9):
                np.testing.assert_allclose(X[t], P[t], rtol=0, atol=1e-15)
                np.testing.assert_allclose(P[t], p0 + a0 * t / 8.0, rtol=0, atol=1e-15)


def test_flow_convergence_orders():
    """Against a fine Heun/improved-Euler solution: the improved-Euler flow on a Heun shot
    converges at second order in dt, the plain Euler flow at first order (a misplaced
    stage of the improved scheme would drop it to first order)."""
    from oracle import deformations as dfo
    rng = np.random.default_rng(21)
    p0 = rng.uniform(size=(12, 3))
    a0 = rng.normal(scale=0.3, size=(12, 3))
    x0 = rng.uniform(size=(20, 3))
    w = 0.4

    def final(T, improved):
        P, A = dfo.shoot(p0, a0, w, T, scheme="heun")
        return dfo.flow(P, A, x0, w, T, improved_euler=improved)[-1]

    ref = final(641, True)
    e = {(T, imp): np.abs(final(T, imp) - ref).max() for T in (11, 21, 41) for imp in (True, False)}
    r_imp = e[(11, True)] / e[(21, True)], e[(21, True)] / e[(41, True)]
    r_eul = e[(11, False)] / e[(21, False)], e[(21, False)] / e[(41, False)]
    assert all(3.3 < r < 4.7 for r in r_imp), r_imp
    assert all(1.6 < r < 2.4 for r in r_eul), r_eul


def test_reference_adjoint_rates_are_jacobian_transposes():
    """The six terms of the reference's ComputeUpdateAt (AdjointEquationsIntegrator.cxx:
    476-567) and its landmark-cotangent rate, as restated in the oracle, are the transposed
    Jacobians of the forward fields, checked group by group against central differences:
    xi' = J_F(z)^T xi + J_z v(x; z)^T theta with F = (K a, -G^T a), v(x; z) = K(x, p) a, and
    theta' = J_x v(x; z)^T theta.  A dropped term, a flipped sign or a transposed operand
    in any group fails one of the comparisons."""
    from oracle import deformations as dfo
    from oracle.oracle import ExactKernel
    rng = np.random.default_rng(7)
    n, m, d, w = 5, 7, 3, 0.8
    P = rng.normal(size=(n, d)) * 0.5
    A = rng.normal(size=(n, d))
    X = rng.normal(size=(m, d)) * 0.6
    th = rng.normal(size=(m, d))
    xp = rng.normal(size=(n, d))
    xm = rng.normal(size=(n, d))
    zero_n, zero_m = np.zeros((n, d)), np.zeros((m, d))

    def K(Y, W):
        k = ExactKernel(d, w)
        k.SetSources(Y)
        k.SetWeights(W)
        return k

    def hamiltonian_pairing(p, a):  # xi . F(p, a)
        k = K(p, a)
        fa = -np.einsum("ikl,ik->il", k.ConvolveGradient(p), a)
        return np.sum(xp * k.Convolve(p)) + np.sum(xm * fa)

    def coupling(p, a, x):  # theta . v(x; p, a)
        return np.sum(th * K(p, a).Convolve(x))

    def fd(fun, Z, h=1e-6):
        g = np.zeros_like(Z)
        for idx in np.ndindex(Z.shape):
            E = np.zeros_like(Z)
            E[idx] = h
            g[idx] = (fun(Z + E) - fun(Z - E)) / (2 * h)
        return g

    # xi part alone (theta = 0): J_F^T xi
    up, um = dfo.reference_update(P, A, X, zero_m, xp, xm, w)
    np.testing.assert_allclose(up, fd(lambda p: hamiltonian_pairing(p, A), P), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(um, fd(lambda a: hamiltonian_pairing(P, a), A), rtol=1e-6, atol=1e-7)
    # landmark coupling alone (xi = 0): J_z v^T theta
    cp_, cm_ = dfo.reference_update(P, A, X, th, zero_n, zero_n, w)
    np.testing.assert_allclose(cp_, fd(lambda p: coupling(p, A, X), P), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(cm_, fd(lambda a: coupling(P, a, X), A), rtol=1e-6, atol=1e-7)
    # both together are the sum (the rate is linear in (xi, theta))
    bp, bm = dfo.reference_update(P, A, X, th, xp, xm, w)
    np.testing.assert_allclose(bp, up + cp_, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(bm, um + cm_, rtol=1e-12, atol=1e-12)
    # theta rate: J_x v^T theta
    np.testing.assert_allclose(dfo.reference_theta_rate(P, A, X, th, w),
                               fd(lambda x: coupling(P, A, x), X), rtol=1e-6, atol=1e-7)


def test_jump_schedule_walks_the_list_as_the_reference_loop():
    """AdjointEquationsIntegrator's jump walk (AEI.cxx:69-131) traced by hand: the index
    moves down the list only under improved Euler; a last jump off T-1 starts theta at
    zero; one condition is no jump; time 0 is never reached."""
    from oracle.deformations import jump_schedule
    assert jump_schedule([0, 2, 4], 5, False) == (2, {2: 1})
    assert jump_schedule([0, 2, 4], 5, True) == (2, {2: 1})
    assert jump_schedule([1, 2, 3, 5], 6, False) == (3, {3: 2})
    assert jump_schedule([1, 2, 3, 5], 6, True) == (3, {3: 2, 2: 1, 1: 0})
    assert jump_schedule([1, 3], 6, False) == (None, {3: 1})
    assert jump_schedule([1, 3], 6, True) == (None, {3: 1, 1: 0})
    assert jump_schedule([2], 6, True) == (0, {})
    assert jump_schedule(None, 6, True) == (0, {})


def test_shoot_vjp_is_the_transpose_jacobian_of_the_shooting_field():
    """oracle shoot_vjp against central differences of mu . F(p, a), F the Euler rates of
    Diffeos::Shoot (Diffeos.cxx:617-632)."""
    from oracle import deformations
    from oracle.deformations import _kernel
    rng = np.random.default_rng(21)
    n, d, width = 12, 3, 0.45
    p, a = rng.uniform(size=(n, d)), rng.normal(size=(n, d))
    mp, ma = rng.normal(size=(n, d)), rng.normal(size=(n, d))

    def objective(pp, aa):
        k = _kernel(d, width, pp, aa)
        fp = k.Convolve(pp)
        fa = -np.einsum("ikl,ik->il", k.ConvolveGradient(pp), aa)
        return float(np.sum(mp * fp) + np.sum(ma * fa))

    jp, ja = deformations.shoot_vjp(p, a, mp, ma, width)
    eps = 1e-6
    for arr, got in ((p, jp), (a, ja)):
        fd = np.zeros_like(arr)
        for i in range(n):
            for k in range(d):
                plus, minus = arr.copy(), arr.copy()
                plus[i, k] += eps
                minus[i, k] -= eps
                args_p = (plus, a) if arr is p else (p, plus)
                args_m = (minus, a) if arr is p else (p, minus)
                fd[i, k] = (objective(*args_p) - objective(*args_m)) / (2 * eps)
        assert np.linalg.norm(got - fd) / np.linalg.norm(fd) < 1e-7
    # a transposed operand or a dropped term fails: swapping the roles of mu_p and a does
    jp_bad, _ = deformations.shoot_vjp(p, mp, a, ma, width)
    assert np.linalg.norm(jp_bad - jp) / np.linalg.norm(jp) > 1e-2
