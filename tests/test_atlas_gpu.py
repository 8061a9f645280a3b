"""GPU deterministic atlas of surfaces (dfm_atlas_*): log-likelihood against the oracle
composition (oracle/atlas.py), gradient against central differences of the GPU
log-likelihood, batched subjects against one-subject atlases, template smoothing, FP32C
against FP64, and HIP-graph replay."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MESH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _problem(seed=0):
    from oracle import surfaces as so
    V, F = so.read_vtk_polydata(os.path.join(MESH, "sphere80.vtk"))
    V2, F2 = so.read_vtk_polydata(os.path.join(MESH, "sphere1280.vtk"))
    rng = np.random.default_rng(seed)
    targets = [
        (V * 1.1 + rng.normal(scale=0.01, size=V.shape), F),
        (V2 * 0.95 + 0.05 * np.sin(3 * V2[:, [1, 2, 0]]), F2),
        (V + np.array([0.05, -0.03, 0.02]) + rng.normal(scale=0.01, size=V.shape), F),
    ]
    g = np.linspace(-0.6, 0.6, 3)
    cp = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    momenta = rng.normal(scale=0.02, size=(len(targets), cp.shape[0], 3))
    return V, F, cp, momenta, targets


PARAMS = dict(deformation_width=0.4, attachment_width=0.2, data_sigma_squared=0.01,
              number_of_time_points=6, checkpoint_every=2)


def _atlas(V, F, cp, targets, kind="varifold", precision="fp64", **kw):
    from deformetrica_code_amd.atlas import HIPDeterministicAtlas
    args = dict(PARAMS)
    args.update(kw)
    return HIPDeterministicAtlas(F, V.shape[0], cp.shape[0], targets, kind=kind,
                                 precision=precision, **args)


@pytest.mark.parametrize("kind", ["current", "varifold"])
def test_loglik_matches_oracle(kind):
    from oracle import atlas as oa
    V, F, cp, mom, targets = _problem(1)
    A = _atlas(V, F, cp, targets, kind)
    (data, reg), res = A.ComputeLogLikelihood(cp, mom, V)
    rd, rr, rres = oa.loglik(cp, mom, V, F, targets, kind=kind, deformation_width=0.4,
                             attachment_width=0.2, data_sigma_squared=0.01, T=6)
    assert abs(data - rd) <= 1e-10 * abs(rd)
    assert abs(reg - rr) <= 1e-10 * abs(rr)
    np.testing.assert_allclose(res, rres, rtol=1e-10)


@pytest.mark.parametrize("kind,scheme", [("varifold", "euler"), ("current", "heun")])
def test_gradient_matches_central_differences(kind, scheme):
    """Directional derivatives of loglik[0] + loglik[1] along random directions in the
    control points, the momenta and the template: the adjoint gradient agrees with central
    differences (eps = 1e-4) to 1e-5 relative."""
    V, F, cp, mom, targets = _problem(2)
    A = _atlas(V, F, cp, targets, kind, scheme=scheme)
    (ll, gcp, gmom, gtpl) = A.ComputeCompleteLogLikelihoodGradient(cp, mom, V)
    f = lambda c, a, v: float(np.sum(A.ComputeLogLikelihood(c, a, v)[0]))
    rng = np.random.default_rng(3)
    eps = 1e-4
    for name, g in (("cp", gcp), ("mom", gmom), ("tpl", gtpl)):
        D = rng.normal(size=g.shape)

        def at(h):
            if name == "cp":
                return cp + h * D, mom, V
            if name == "mom":
                return cp, mom + h * D, V
            return cp, mom, V + h * D

        fd1 = (f(*at(eps)) - f(*at(-eps))) / (2 * eps)
        fd2 = (f(*at(eps / 2)) - f(*at(-eps / 2))) / eps
        fd = (4 * fd2 - fd1) / 3    # Richardson: the O(eps^2) truncation cancels
        an = float(np.sum(g * D))
        print(name, an, fd1, fd, abs(an - fd1) / abs(fd1), abs(an - fd) / abs(fd))
        assert abs(an - fd1) <= 5e-5 * abs(fd1), (name, an, fd1)
        assert abs(an - fd) <= 1e-6 * abs(fd), (name, an, fd)


def test_batched_subjects_equal_one_subject_atlases():
    V, F, cp, mom, targets = _problem(4)
    A = _atlas(V, F, cp, targets)
    ll, gcp, gmom, gtpl = A.ComputeCompleteLogLikelihoodGradient(cp, mom, V)
    acc_ll, acc_cp, acc_tpl = np.zeros(2), np.zeros_like(gcp), np.zeros_like(gtpl)
    for s, t in enumerate(targets):
        B = _atlas(V, F, cp, [t])
        l1, c1, m1, t1 = B.ComputeCompleteLogLikelihoodGradient(cp, mom[s:s + 1], V)
        acc_ll += l1
        acc_cp += c1
        acc_tpl += t1
        assert rel(gmom[s], m1[0]) < 1e-12
    assert rel(ll, acc_ll) < 1e-12
    assert rel(gcp, acc_cp) < 1e-12
    assert rel(gtpl, acc_tpl) < 1e-12


def test_template_gradient_smoothing_is_the_reference_convolution():
    from oracle import atlas as oa
    V, F, cp, mom, targets = _problem(5)
    plain = _atlas(V, F, cp, targets).ComputeCompleteLogLikelihoodGradient(cp, mom, V)[3]
    smooth = _atlas(V, F, cp, targets, smoothing_width=0.3).ComputeCompleteLogLikelihoodGradient(
        cp, mom, V)[3]
    assert rel(smooth, oa.smooth_template_gradient(V, plain, 0.3)) < 1e-12


def test_fp32c_atlas_gradient_close_to_fp64():
    V, F, cp, mom, targets = _problem(6)
    out = {p: _atlas(V, F, cp, targets, precision=p).ComputeCompleteLogLikelihoodGradient(cp, mom, V)
           for p in ("fp64", "fp32c")}
    for k in range(4):
        assert rel(out["fp32c"][k], out["fp64"][k]) < 1e-5, k


def test_graph_replay_of_the_atlas_gradient():
    from deformetrica_code_amd.device import DeviceArray
    V, F, cp, mom, targets = _problem(7)
    A = _atlas(V, F, cp, targets, precision="fp32c")
    st = A.stream
    dcp, dmom, dV = (DeviceArray.from_numpy(v, st) for v in (cp, mom, V))

    def read():
        st.synchronize()
        return [a.numpy(st) for a in (A.loglik, A.grad_cp, A.grad_momenta, A.grad_template)]

    g = A.capture(dcp, dmom, dV)
    rng = np.random.default_rng(8)
    for r in range(3):
        m = mom * (1.0 + 0.2 * r) + rng.normal(scale=1e-3, size=mom.shape)
        dmom.upload(m, st)
        A.gradient_device(dcp, dmom, dV)
        ref = read()
        for a in (A.loglik, A.grad_cp, A.grad_momenta, A.grad_template):
            a.upload(np.zeros(a.shape), st)
        g.launch()
        out = read()
        for u, v in zip(out, ref):
            np.testing.assert_array_equal(u, v)


def _polyline_problem(d, seed):
    """An open template curve and three ragged polyline targets in dimension d."""
    rng = np.random.default_rng(seed)
    s = np.linspace(0, 1.5 * np.pi, 41)
    V = np.zeros((41, d))
    V[:, 0], V[:, 1] = np.cos(s), 0.6 * np.sin(s)
    if d == 3:
        V[:, 2] = 0.2 * s
    E = np.stack([np.arange(40), np.arange(1, 41)], axis=1)
    targets = []
    for k, n in enumerate((30, 52, 41)):
        t = np.linspace(0, 1.5 * np.pi, n)
        W = np.zeros((n, d))
        W[:, 0], W[:, 1] = (1.05 + 0.05 * k) * np.cos(t), (0.55 + 0.04 * k) * np.sin(t + 0.1 * k)
        if d == 3:
            W[:, 2] = 0.2 * t + 0.05 * k
        W += rng.normal(scale=0.005, size=W.shape)
        F = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1)
        if k == 1:
            F = F[::-1, ::-1].copy()           # reversed order and orientation
        targets.append((W, F))
    g = np.linspace(-1.0, 1.0, 4)
    grids = np.meshgrid(*([g] * d), indexing="ij")
    cp = np.stack(grids, -1).reshape(-1, d)
    momenta = rng.normal(scale=0.03, size=(len(targets), cp.shape[0], d))
    return V, E, cp, momenta, targets


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("kind", ["current", "varifold"])
def test_polyline_atlas_matches_the_oracle(d, kind):
    """The batched atlas with polyline subjects (segments, dimension 2 or 3) against the
    oracle's one-object atlas with the reference's continuous adjoint: log-likelihood,
    residuals and every gradient block to 1e-10; template smoothing included."""
    from oracle import atlas as oa
    V, E, cp, mom, targets = _polyline_problem(d, 10 + d)
    print(f"polyline atlas d={d} kind={kind}", flush=True)
    common = dict(deformation_width=0.6, T=6, improved_euler=False)
    A = _atlas(V, E, cp, targets, kind, deformation_width=0.6, attachment_width=0.25,
               data_sigma_squared=0.05, improved_euler=False, adjoint="continuous",
               smoothing_width=0.3)
    (data, reg), res = A.ComputeLogLikelihood(cp, mom, V)
    objects = [(kind, E, 0.25, 0.05)]
    subj = [[t] for t in targets]
    rd, rr, rres = oa.multi_loglik(cp, mom, [V], objects, subj, **common)
    assert abs(data - rd) <= 1e-10 * abs(rd)
    assert abs(reg - rr) <= 1e-10 * abs(rr)
    np.testing.assert_allclose(res, rres[:, 0], rtol=1e-10)
    ll, gcp, gmom, gtpl = A.ComputeCompleteLogLikelihoodGradient(cp, mom, V)
    rcp, rmom, sob, _ = oa.multi_gradient_reference(cp, mom, [V], objects, subj,
                                                    smoothing_width=0.3, **common)
    assert rel(gcp, rcp) < 1e-10
    assert rel(gmom, rmom) < 1e-10
    assert rel(gtpl, sob[0]) < 1e-10


def test_polyline_atlas_fp32c_against_fp64():
    V, E, cp, mom, targets = _polyline_problem(2, 20)
    out = {}
    for prec in ("fp64", "fp32c"):
        A = _atlas(V, E, cp, targets, "varifold", precision=prec, deformation_width=0.6,
                   attachment_width=0.25, data_sigma_squared=0.05)
        out[prec] = A.ComputeCompleteLogLikelihoodGradient(cp, mom, V)
    for a, b in zip(out["fp32c"], out["fp64"]):
        assert rel(a, b) < 1e-5


def test_cells_and_dimension_are_checked():
    from deformetrica_code_amd import _lib
    V, E, cp, mom, targets = _polyline_problem(2, 21)
    with pytest.raises(_lib.DfmError, match="outside"):
        bad = [(targets[0][0], targets[0][1] + 100)] + targets[1:]
        _atlas(V, E, cp, bad, "current")
    Vs, Fs, cps, moms, tgs = _problem(5)
    flat = [(v[:, :2], f) for v, f in tgs]       # triangles in 2D
    with pytest.raises(_lib.DfmError, match="triangles need 3"):
        _atlas(Vs[:, :2], Fs, cps[:, :2], flat, "current")


@pytest.mark.parametrize("d", [2, 3])
def test_graph_replay_of_the_polyline_atlas_gradient(d):
    """The polyline atlas gradient captured once and replayed on new momenta, with eager
    calls in between: every replay equals the eager evaluation bit for bit."""
    from deformetrica_code_amd.device import DeviceArray
    V, E, cp, mom, targets = _polyline_problem(d, 30 + d)
    A = _atlas(V, E, cp, targets, "varifold", precision="fp32c", deformation_width=0.6,
               attachment_width=0.25, data_sigma_squared=0.05, smoothing_width=0.3)
    st = A.stream
    dcp, dmom, dV = (DeviceArray.from_numpy(v, st) for v in (cp, mom, V))

    def read():
        st.synchronize()
        return [a.numpy(st) for a in (A.loglik, A.grad_cp, A.grad_momenta, A.grad_template)]

    g = A.capture(dcp, dmom, dV)
    rng = np.random.default_rng(9)
    for r in range(3):
        dmom.upload(mom * (1.0 + 0.3 * r) + rng.normal(scale=1e-3, size=mom.shape), st)
        A.gradient_device(dcp, dmom, dV)
        ref = read()
        for a in (A.loglik, A.grad_cp, A.grad_momenta, A.grad_template):
            a.upload(np.zeros(a.shape), st)
        g.launch()
        for u, v in zip(read(), ref):
            np.testing.assert_array_equal(u, v)
