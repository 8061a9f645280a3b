"""ORACLE — test infrastructure only.

FP64 CPU restatement of Deformetrica's deformation integrators that call the kernel-sum
engine, on top of the oracle ExactKernel (oracle/exact_kernel.c):

  shoot            Diffeos::Shoot — explicit Euler (Diffeos.cxx:589-659), with the
                   commented-out Heun variant (Diffeos.cxx:634-652) as scheme="heun"
  flow             Diffeos::FlowLandmarkPointsTrajectory (Diffeos.cxx:661-710)
  shoot_vjp        the shooting field's transpose Jacobian, from the reference's kernel ops
  parallel_transport  Diffeos::ParallelTransport (Diffeos.cxx:293-582)
  adjoint_reference   AdjointEquationsIntegrator (AdjointEquationsIntegrator.cxx:33-136,
                   414-567): the reference's continuous adjoint, discretised with Heun

Pinned by tests/test_deformations.py: the reference's own parallel-transport golden
vectors (tests/unit_tests/parallel-transport/data, GroundTruthMom.txt at 1e-4), the
Hamiltonian's conservation to O(dt), the adjoint's rates (reference_update,
reference_theta_rate) term group by term group against central differences of the forward
fields they transpose, and central differences of the shooting objective against the
reference adjoint (to O(dt), the order at which a continuous adjoint discretised
independently of the forward scheme agrees with the discrete gradient).
Matrices are numpy float64 (rows x cols).
"""
from __future__ import annotations

import numpy as np

from .oracle import ExactKernel


def _kernel(dim, width, Y, W):
    k = ExactKernel(dim, width)
    k.SetSources(Y)
    k.SetWeights(W)
    return k


def time_step(T, t0=0.0, tN=1.0):
    return (tN - t0) / (T - 1)


def shoot(p0, a0, width, T, t0=0.0, tN=1.0, scheme="euler"):
    """Control-point and momenta trajectories, lists of T (n x d) arrays."""
    p0 = np.asarray(p0, dtype=np.float64)
    a0 = np.asarray(a0, dtype=np.float64)
    n, d = p0.shape
    P = [p0.copy() for _ in range(T)]
    A = [np.zeros((n, d)) for _ in range(T)]
    A[0] = a0.copy()
    if np.linalg.norm(a0) < 1e-20:  # Diffeos.cxx:609-611
        return P, A
    dt = time_step(T, t0, tN)

    def field(p, a):
        k = _kernel(d, width, p, a)
        G = k.ConvolveGradient(p)           # (n, d, d) varlin, Diffeos.cxx:624
        dpos = k.Convolve(p)                # Diffeos.cxx:626
        dmom = np.einsum("ikl,ik->il", G, a)  # G_i^T a_i, Diffeos.cxx:628-629
        return dpos, dmom

    for t in range(T - 1):
        dpos, dmom = field(P[t], A[t])
        P[t + 1] = P[t] + dpos * dt
        A[t + 1] = A[t] - dmom * dt
        if scheme == "heun":
            dpos2, dmom2 = field(P[t + 1], A[t + 1])
            P[t + 1] = P[t] + (dpos + dpos2) * (dt * 0.5)
            A[t + 1] = A[t] - (dmom + dmom2) * (dt * 0.5)
    return P, A


def shoot_vjp(p, a, mu_p, mu_a, width):
    """Transpose Jacobian of the shooting field F(p, a) = (K(p,p) a, -G(p,a)^T a) (the
    rates of Diffeos::Shoot, Diffeos.cxx:617-632) applied to (mu_p, mu_a), composed from
    the kernel operations the reference's adjoint uses for the same terms
    (AdjointEquationsIntegrator.cxx:476-567):
      d/dp = ConvolveGradient(p, mu_p) [weights a] + ConvolveGradient(p, a) [weights mu_p]
             + ConvolveSpecialHessian(mu_a) [weights a]
      d/da = Convolve(p) [weights mu_p] - (G(p,a) mu_a - sum_l ConvolveGradient(p, l)
             [weights a * mu_a_l])
    Pinned against central differences of F in tests/test_deformations.py."""
    p, a, mu_p, mu_a = (np.asarray(v, dtype=np.float64) for v in (p, a, mu_p, mu_a))
    n, d = p.shape
    ka = _kernel(d, width, p, a)
    kmp = _kernel(d, width, p, mu_p)
    jp = (ka.ConvolveGradient(p, mu_p) + kmp.ConvolveGradient(p, a)
          + ka.ConvolveSpecialHessianGradDiff(mu_a))
    G = ka.ConvolveGradient(p)  # (n, d, d): G[i, k, l] = sum_j a_j[k] d/dp_l K
    diff = np.einsum("ikl,il->ik", G, mu_a)
    for l in range(d):
        diff -= _kernel(d, width, p, a * mu_a[:, l:l + 1]).ConvolveGradient(p, l)
    ja = kmp.Convolve(p) - diff
    return jp, ja


def flow(P, A, x0, width, T, t0=0.0, tN=1.0, improved_euler=True):
    """Landmark trajectory (list of T (m x d) arrays) in the flow of (P, A)."""
    x0 = np.asarray(x0, dtype=np.float64)
    d = x0.shape[1]
    X = [x0.copy() for _ in range(T)]
    if np.linalg.norm(A[0]) < 1e-20:  # Diffeos.cxx:684-685
        return X
    dt = time_step(T, t0, tN)
    for t in range(T - 1):
        v = _kernel(d, width, P[t], A[t]).Convolve(X[t])
        X[t + 1] = X[t] + v * dt
        if improved_euler:
            v2 = _kernel(d, width, P[t + 1], A[t + 1]).Convolve(X[t + 1])
            X[t + 1] = X[t] + (v + v2) * (dt * 0.5)
    return X


def parallel_transport(P, A, width, T, initial_momenta, initial_cp, initial_time, target_times,
                       t0=0.0, tN=1.0):
    """Diffeos::ParallelTransport (Diffeos.cxx:293-582): momenta transported along the
    geodesic (P, A) at the target times."""
    d = P[0].shape[1]
    nb = len(target_times)
    if (np.sum(A[0] ** 2) < 1e-20 or np.sum(initial_momenta ** 2) < 1e-20 or nb == 0 or T <= 1):
        return [initial_momenta.copy() for _ in range(nb)]
    h = (tN - t0) / (T - 1)
    eps = h

    def to_index(t):
        c = (t - t0) / h
        i = int(c)
        return i + 1 if (c - i) >= 0.5 else i

    initial_index = to_index(initial_time)
    target_indices = [to_index(t) for t in target_times]
    final_index = target_indices[-1]
    n_steps = final_index - initial_index + 1
    if final_index == initial_index:
        return [initial_momenta.copy() for _ in range(nb)]

    def conv(Y, W, X):
        return _kernel(d, width, Y, W).Convolve(X)

    def dot(a, b):
        return float(np.sum(a * b))

    initial_velocity = conv(initial_cp, initial_momenta, P[0])
    velocities = [initial_velocity.copy() for _ in range(n_steps)]
    transport = [None] * n_steps
    init_wv = init_w2 = 0.0
    wv = dot(initial_momenta, conv(P[initial_index], A[initial_index], initial_cp))
    w2 = dot(initial_momenta, conv(initial_cp, initial_momenta, initial_cp))
    v2 = 0.0
    alpha, beta = 1.0, 0.0
    index = 0
    for t in range(initial_index, final_index):
        k = ExactKernel(d, width)
        k.SetSources(P[t])
        k.SetWeights(A[t])
        dpos = k.Convolve(P[t])
        kinv = k.ConvolveInverse(velocities[index])
        if t == initial_index:
            init_wv = dot(kinv, k.Convolve(P[t]))
            init_w2 = dot(kinv, conv(P[t], kinv, P[t]))
        if t > initial_index:
            q = conv(P[t], kinv, P[t])
            prop_n2 = dot(kinv, q)
            prop_wv = dot(A[t], q)
            alpha = np.sqrt((init_w2 * v2 - init_wv * wv) / (prop_n2 * v2 - prop_wv * prop_wv))
            beta = (init_wv - alpha * prop_wv) / v2
        transport[index] = alpha * kinv + beta * A[t]
        q = conv(P[t], transport[index], P[t])
        wv = dot(A[t], q)
        w2 = dot(kinv, q)
        v2 = dot(A[t], conv(P[t], A[t], P[t]))
        mom_pos = A[t] + eps * kinv
        mom_neg = A[t] - eps * kinv

        def dmom(mom):
            G = _kernel(d, width, P[t], mom).ConvolveGradient(P[t])
            return np.einsum("ikl,ik->il", G, mom)

        dm_pos, dm_neg = dmom(mom_pos), dmom(mom_neg)
        cp1 = P[t] + h / 2 * (dpos + eps * velocities[index])
        mom1 = mom_pos - h / 2 * dm_pos
        cp2 = P[t] + h * conv(cp1, mom1, cp1)
        plus = cp2 * (1 / (2 * eps))
        cp1 = P[t] + h / 2 * (dpos - eps * velocities[index])
        mom1 = mom_neg - h / 2 * dm_neg
        cp2 = P[t] + h * conv(cp1, mom1, cp1)
        minus = cp2 * (1 / (2 * eps))
        velocities[index + 1] = (plus - minus) / eps
        index += 1
    k = ExactKernel(d, width)
    k.SetSources(P[final_index - 1])
    k.SetWeights(A[final_index - 1])
    transport[index] = k.ConvolveInverse(velocities[index])
    # the reference indexes its per-step buffer by the absolute time index t
    # (parallelTransport[t], Diffeos.cxx:571-579), which is the step t - initial_index
    # when the initial time is t0; the step index is used here
    out = []
    oi = 0
    for t in range(initial_index, final_index + 1):
        if oi < nb and t == target_indices[oi]:
            out.append(transport[t - initial_index])
            oi += 1
    return out


def hamiltonian(p, a, width):
    """H = 1/2 a^T K(p, p) a, conserved along the geodesic."""
    return 0.5 * float(np.sum(a * _kernel(p.shape[1], width, p, a).Convolve(p)))


def reference_theta_rate(P_s, A_s, X_s, theta_s, width):
    """Landmark-cotangent rate of AdjointEquationsIntegrator (AEI.cxx:96-133): the kernel of
    (P_s, A_s) differentiated at the landmarks, contracted with theta."""
    d = P_s.shape[1]
    return _kernel(d, width, P_s, A_s).ConvolveGradient(X_s, theta_s)


def reference_update(P_s, A_s, X_s, theta_s, xp_s, xm_s, width):
    """ComputeUpdateAt (AEI.cxx:476-567), landmark branch: the rates (d xi_pos, d xi_mom)
    at state s, the sum of its six kernel terms."""
    n, d = P_s.shape
    eta = _kernel(d, width, X_s, theta_s)
    dxi1 = eta.ConvolveGradient(P_s, A_s)
    dxi2 = eta.Convolve(P_s)
    axp = np.concatenate([A_s, xp_s], axis=1)
    k = _kernel(d, width, P_s, axp)
    kaxp = k.Convolve(P_s)
    g = k.ConvolveGradient(P_s)  # (n, 2d, d)
    dxi3 = np.einsum("ikl,ik->il", g[:, :d, :], xp_s) + np.einsum("ikl,ik->il", g[:, d:, :], A_s)
    dxi5 = kaxp[:, d:]
    dxi6 = np.zeros((n, d))
    for dim in range(d):
        Wd = xm_s[:, dim:dim + 1] * A_s
        dxi6 += _kernel(d, width, P_s, Wd).ConvolveGradient(P_s, dim)
    dxi6 -= np.einsum("ikl,il->ik", g[:, :d, :], xm_s)
    dxi4 = _kernel(d, width, P_s, A_s).ConvolveSpecialHessian(xm_s)
    return dxi1 + dxi3 + dxi4, dxi2 + dxi5 + dxi6


def jump_schedule(jump_times, T, improved_euler):
    """Where AdjointEquationsIntegrator::IntegrateAdjointOfLandmarkPointsEquations
    (AEI.cxx:59-136) uses the conditions of a list with jump times: (index of the condition
    that starts theta at T-1, or None for zero; {step t: condition index}), the condition
    of step t being added to the rate of theta over t -> t-1.  Its index walks down the
    list only in the improved-Euler branch, so under standard Euler at most one jump after
    the terminal one is applied; a list of one condition has no jumps and starts theta
    whatever its time; a jump at time 0 is never reached."""
    if jump_times is None or len(jump_times) <= 1:
        return 0, {}
    subj = len(jump_times) - 1
    terminal = None
    if jump_times[subj] == T - 1:
        terminal = subj
        subj = max(subj - 1, 0)
    steps = {}
    for t in range(T - 1, 0, -1):
        if t == jump_times[subj]:
            steps[t] = subj
            if improved_euler:
                subj = max(subj - 1, 0)
    return terminal, steps


def adjoint_reference(P, A, X, theta_T, width, T, t0=0.0, tN=1.0, improved_euler=True,
                      jump_times=None):
    """AdjointEquationsIntegrator with landmark points only (AEI.cxx:59-136, 414-567):
    returns (xi_pos_0, xi_mom_0, theta_0), the reference's gradient of E(x_T) w.r.t. the
    initial control points, momenta and landmarks for dE/dx_T = theta_T.  With jump_times
    (Diffeos::IntegrateAdjointEquations with a list, Diffeos.cxx:1393-1443), theta_T is
    the list of conditions at those time indices, used as ``jump_schedule`` says."""
    d = P[0].shape[1]
    n = P[0].shape[0]
    dt = time_step(T, t0, tN)
    if jump_times is None:
        conditions = [np.asarray(theta_T, dtype=np.float64)]
    else:
        conditions = [np.asarray(c, dtype=np.float64) for c in theta_T]
    terminal, steps = jump_schedule(jump_times, T, improved_euler)
    theta = [None] * T
    theta[T - 1] = (conditions[terminal].copy() if terminal is not None
                    else np.zeros_like(conditions[0]))
    for t in range(T - 1, 0, -1):  # AEI.cxx:96-133
        dth = reference_theta_rate(P[t], A[t], X[t], theta[t], width)
        if t in steps:
            dth = dth + conditions[steps[t]]
        theta[t - 1] = theta[t] + dth * dt
        if improved_euler:
            dth2 = reference_theta_rate(P[t - 1], A[t - 1], X[t - 1], theta[t - 1], width)
            if t in steps:
                dth2 = dth2 + conditions[steps[t]]
            theta[t - 1] = theta[t] + (dth + dth2) * (dt * 0.5)
    xp = [np.zeros((n, d)) for _ in range(T)]
    xm = [np.zeros((n, d)) for _ in range(T)]

    def update(s):
        return reference_update(P[s], A[s], X[s], theta[s], xp[s], xm[s], width)

    for t in range(T - 1, 0, -1):  # AEI.cxx:438-462
        dpos, dmom = update(t)
        xp[t - 1] = xp[t] + dpos * dt
        xm[t - 1] = xm[t] + dmom * dt
        if improved_euler:
            dpos2, dmom2 = update(t - 1)
            xp[t - 1] = xp[t] + (dpos + dpos2) * (dt * 0.5)
            xm[t - 1] = xm[t] + (dmom + dmom2) * (dt * 0.5)
    return xp[0], xm[0], theta[0]


def read_dlm(path):
    """Deformetrica's DLM text matrices (src/io/MatrixDLM.cxx): whitespace-separated rows;
    the multiple-matrix form starts with "count rows cols"."""
    rows = [line.split() for line in open(path) if line.strip()]
    first = rows[0]
    if len(first) == 3 and len(rows) > 1 and len(rows[1]) != 3 and float(first[0]).is_integer():
        cnt, r, c = (int(float(v)) for v in first)
        data = np.array([[float(v) for v in row] for row in rows[1:1 + r]])
        assert data.shape == (r, c)
        return data
    return np.array([[float(v) for v in row] for row in rows])
