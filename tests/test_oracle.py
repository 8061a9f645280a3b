"""Pins the FP64 oracle (oracle/exact_kernel.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip in the restatement fails it: sigma^2 vs
2 sigma^2 or sigma in the exponent (one source at a known distance), a wrong sign or index
in the gradients (central differences of the oracle's own Convolve), a transposed varlin
block, a wrong Hessian term (differences of the pinned gradient), and the two forms of
ConvolveSpecialHessian that the reference itself tests against each other
(tests/unit_tests/kernels/TestKernelPrecisionCUDA.cxx:97-106).
"""
import numpy as np
import pytest


def test_mt19937_known_answers(oracle_mod):
    # C++ [rand.predef]: the 10000th output of a default-constructed mt19937 is 4123659995.
    assert oracle_mod.mt19937_stream(5489, 10000)[-1] == 4123659995
    # std::mt19937(1)() == 1791095845
    assert oracle_mod.mt19937_stream(1, 1)[0] == 1791095845


def test_reference_fixture_generator(oracle_mod):
    # AbstractTestKernelPrecision.cxx:21-29 reseeds with 1 for every matrix, filling row by row.
    X = oracle_mod.reference_fixture(501, 2)
    Y = oracle_mod.reference_fixture(201, 2)
    assert np.array_equal(X[:201], Y)
    s = oracle_mod.mt19937_stream(1, 9).astype(np.float64)
    Z = oracle_mod.reference_fixture(3, 3)
    assert np.array_equal(Z.ravel(), np.exp(-s / 4294967295.0))
    assert X.min() > np.exp(-1.0) - 1e-15 and X.max() <= 1.0


def test_convolve_single_source_closed_form(oracle_mod):
    h = 0.37
    k = oracle_mod.ExactKernel(3, h)
    y = np.array([[0.2, -0.1, 0.4]])
    k.SetSources(y)
    k.SetWeights(np.array([[1.5, -2.0, 0.25]]))
    # a target exactly one kernel width away: K = exp(-1) (sigma^2 in the denominator, no 2)
    u = np.array([3.0, -1.0, 2.0]) / np.sqrt(14.0)
    X = y + h * u[None, :]
    V = k.Convolve(X)
    np.testing.assert_allclose(V, np.exp(-1.0) * np.array([[1.5, -2.0, 0.25]]), rtol=1e-15)
    # coincident point: K = 1
    np.testing.assert_allclose(k.Convolve(y), [[1.5, -2.0, 0.25]], rtol=0, atol=0)


def test_convolve_wide_kernel_limit_and_width_clamp(oracle_mod):
    rng = np.random.default_rng(3)
    Y = rng.uniform(size=(37, 2))
    W = rng.normal(size=(37, 4))
    X = rng.uniform(size=(11, 2))
    k = oracle_mod.ExactKernel(2, 1e9)
    k.SetSources(Y)
    k.SetWeights(W)
    np.testing.assert_allclose(k.Convolve(X), np.tile(W.sum(0), (11, 1)), rtol=1e-12, atol=1e-12)
    # width 0 is clamped to 1e-10 (AbstractKernel.h:85-91): only coincident points survive
    k0 = oracle_mod.ExactKernel(2, 0.0)
    assert k0.GetKernelWidth() == 1e-10
    k0.SetSources(Y)
    k0.SetWeights(W)
    np.testing.assert_array_equal(k0.Convolve(Y), W)


def test_kernel_symmetry(oracle_mod):
    rng = np.random.default_rng(4)
    A = rng.uniform(size=(9, 3))
    B = rng.uniform(size=(13, 3))
    k = oracle_mod.ExactKernel(3, 0.3)
    KAB = np.empty((9, 13))
    for j in range(13):
        k.SetSources(B[j:j + 1])
        k.SetWeights(np.ones((1, 1)))
        KAB[:, j] = k.Convolve(A)[:, 0]
    KBA = np.empty((13, 9))
    for j in range(9):
        k.SetSources(A[j:j + 1])
        k.SetWeights(np.ones((1, 1)))
        KBA[:, j] = k.Convolve(B)[:, 0]
    np.testing.assert_allclose(KAB, KBA.T, rtol=1e-15)


def _fd_conv(k, X, i, l, eps):
    Xp, Xm = X.copy(), X.copy()
    Xp[i, l] += eps
    Xm[i, l] -= eps
    return (k.Convolve(Xp[i:i + 1]) - k.Convolve(Xm[i:i + 1]))[0] / (2 * eps)


@pytest.mark.parametrize("d,dv", [(2, 2), (3, 3), (3, 6), (2, 4)])
def test_gradients_match_central_differences(oracle_mod, d, dv):
    rng = np.random.default_rng(10 + d * 7 + dv)
    h = 0.45
    Y = rng.uniform(size=(23, d))
    W = rng.normal(size=(23, dv))
    X = rng.uniform(size=(7, d))
    alpha = rng.normal(size=(7, dv))
    k = oracle_mod.ExactKernel(d, h)
    k.SetSources(Y)
    k.SetWeights(W)
    G = k.ConvolveGradient(X)  # (n, dv, d)
    g1 = k.ConvolveGradient(X, alpha)
    gd = [k.ConvolveGradient(X, dim) for dim in range(d)]
    eps = 1e-6
    for i in range(X.shape[0]):
        for l in range(d):
            fd = _fd_conv(k, X, i, l, eps)  # d Conv_i / d x_il, (dv,)
            np.testing.assert_allclose(G[i, :, l], fd, rtol=1e-7, atol=1e-8)
            np.testing.assert_allclose(gd[l][i], fd, rtol=1e-7, atol=1e-8)
            np.testing.assert_allclose(g1[i, l], fd @ alpha[i], rtol=1e-7, atol=1e-8)
    with pytest.raises(RuntimeError):
        k.ConvolveGradient(X, d)


@pytest.mark.parametrize("d", [2, 3])
def test_hessian_matches_differences_of_gradient(oracle_mod, d):
    rng = np.random.default_rng(20 + d)
    Y = rng.uniform(size=(17, d))
    W = rng.normal(size=(17, d))
    X = rng.uniform(size=(5, d))
    k = oracle_mod.ExactKernel(d, 0.5)
    k.SetSources(Y)
    k.SetWeights(W)
    H = k.ConvolveHessian(X)  # (n, dv, d, d)
    eps = 1e-6
    for i in range(X.shape[0]):
        for b in range(d):
            Xp, Xm = X.copy(), X.copy()
            Xp[i, b] += eps
            Xm[i, b] -= eps
            fd = (k.ConvolveGradient(Xp[i:i + 1]) - k.ConvolveGradient(Xm[i:i + 1]))[0] / (2 * eps)
            np.testing.assert_allclose(H[i, :, :, b], fd, rtol=1e-6, atol=1e-7)
    for r in range(d):
        for c in range(d):
            np.testing.assert_allclose(k.ConvolveHessian(X, r, c), H[:, :, r, c], rtol=1e-13,
                                       atol=1e-13)


@pytest.mark.parametrize("d", [2, 3])
def test_special_hessian_two_reference_forms_and_brute_force(oracle_mod, d):
    rng = np.random.default_rng(30 + d)
    n = 19
    Y = rng.uniform(size=(n, d))
    beta = rng.normal(size=(n, d))
    xi = rng.normal(size=(n, d))
    h = 0.4
    k = oracle_mod.ExactKernel(d, h)
    k.SetSources(Y)
    k.SetWeights(beta)
    generic = k.ConvolveSpecialHessian(xi)
    closed = k.ConvolveSpecialHessianGradDiff(xi)
    np.testing.assert_allclose(generic, closed, rtol=1e-11, atol=1e-12)
    np.testing.assert_array_equal(k.GetWeights(), beta)  # weights restored (AbstractKernel.cxx:84)
    # brute force: gamma_i = sum_j (b_i.b_j) H(y_i, y_j)(xi_j - xi_i), with H from central
    # differences of the pinned single-source gradient.
    one = oracle_mod.ExactKernel(d, h)
    eps = 1e-6
    brute = np.zeros((n, d))
    for i in range(n):
        for j in range(n):
            one.SetSources(Y[j:j + 1])
            one.SetWeights(np.ones((1, 1)))
            Hij = np.empty((d, d))
            for b in range(d):
                p, m = Y[i:i + 1].copy(), Y[i:i + 1].copy()
                p[0, b] += eps
                m[0, b] -= eps
                Hij[:, b] = (one.ConvolveGradient(p)[0, 0] - one.ConvolveGradient(m)[0, 0]) / (2 * eps)
            brute[i] += (beta[i] @ beta[j]) * Hij @ (xi[j] - xi[i])
    np.testing.assert_allclose(generic, brute, rtol=1e-6, atol=1e-6)


def test_convolve_inverse_round_trip(oracle_mod):
    rng = np.random.default_rng(40)
    Y = rng.uniform(size=(30, 2)) * 10
    X = rng.normal(size=(30, 2))
    k = oracle_mod.ExactKernel(2, 1.5)
    k.SetSources(Y)
    V = k.ConvolveInverse(X)
    k.SetWeights(V)
    np.testing.assert_allclose(k.Convolve(Y), X, rtol=1e-8, atol=1e-9)
    # the reference's solve() is LAPACK gesv (LU, partial pivoting), as numpy.linalg.solve
    K = k.ComputeKernelMatrix(Y)
    np.testing.assert_allclose(V, np.linalg.solve(K, X), rtol=1e-12, atol=1e-12)


def test_threaded_matches_serial(oracle_mod):
    rng = np.random.default_rng(50)
    X = rng.uniform(size=(101, 3))
    A = rng.normal(size=(101, 3))
    Y = rng.uniform(size=(77, 3))
    W = rng.normal(size=(77, 3))
    c, g = oracle_mod.convolve_grad1_threaded(X, A, Y, W, 0.2, 4)
    k = oracle_mod.ExactKernel(3, 0.2)
    k.SetSources(Y)
    k.SetWeights(W)
    np.testing.assert_array_equal(c, k.Convolve(X))
    np.testing.assert_array_equal(g, k.ConvolveGradient(X, A))


def test_row_parallel_sums_match_one_thread(oracle_mod):
    """Sums split over host threads by target rows equal the one-thread sums bit for bit
    (each row's sum is the same C loop whichever thread runs it)."""
    rng = np.random.default_rng(51)
    X = rng.uniform(size=(3000, 3))
    A = rng.normal(size=(3000, 3))
    Y = rng.uniform(size=(2000, 3))
    W = rng.normal(size=(2000, 6))
    k = oracle_mod.ExactKernel(3, 0.2)
    k.SetSources(Y)
    k.SetWeights(W)
    out = []
    for t in (1, 6):
        oracle_mod.set_threads(t)
        out.append((k.Convolve(X), k.ConvolveGradient(X), k.ConvolveGradient(X, A[:, :3].repeat(2, 1)),
                    k.ConvolveGradient(X, 2)))
    oracle_mod.set_threads(16)
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    with pytest.raises(RuntimeError, match="out of bounds"):
        k.ConvolveGradient(X, 3)


def test_mismatch_raises(oracle_mod):
    k = oracle_mod.ExactKernel(3, 1.0)
    k.SetSources(np.zeros((4, 3)))
    k.SetWeights(np.zeros((5, 3)))
    with pytest.raises(RuntimeError, match="mismatch"):
        k.Convolve(np.zeros((2, 3)))
