"""Two of the reference's functional tests run end to end on the GPU, through the
estimators (deformetrica_code_amd.estimators) over the batched atlas (HIPAtlasModel) with
the reference's continuous adjoint, standard Euler and its template-update semantics, in
FP64:

* the banana atlas (configure.ini Test2, GradientAscent, 3 subjects, template updates):
  the reference's own comparison of the final snapshot at its 1e-5 tolerance, and at
  1e-10; every field within 1e-9 relative of the oracle replay;
* the amygdala registration (Test6, FastGradientAscent): every snapshot field within 1e-9
  relative of the oracle replay, and the stored snapshot's discrete path exactly with every
  value within the reference's 1e-8 plus 1e-13 of its magnitude;
* the GPU log-likelihood at the stored amygdala state: the regularity term to 1e-12 of the
  stored one, the data term of the undeformed template to 1e-12 of the stored one;
* FP32C kernel sums take the same optimizer path as FP64;
* the skulls registration (Test5, a varifold polyline) and the parabola atlas (Test8, two
  polyline currents) through the multi-object model (HIPMultiObjectAtlasModel): within
  1e-9 relative of the oracle replays, the stored snapshots' discrete paths exactly and
  every value within the tests' tolerances; the amygdala registration through the same
  model agrees with the batched atlas's replay;
* the skulls regression (Test7) through HIPRegressionModel, whose backward pass is the
  continuous adjoint with the observations as jump conditions: within 1e-9 relative of the
  oracle replay, the snapshot's path exactly, every value within the test's 1e-8.
"""
import numpy as np
import pytest

import functional_cases as fc

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _model(V0, F, targets, params, fixed, *, freeze_template, smoothing, precision="fp64"):
    from deformetrica_code_amd.atlas import HIPAtlasModel, HIPDeterministicAtlas
    atlas = HIPDeterministicAtlas(F, V0.shape[0], fixed["ControlPoints"].shape[0], targets,
                                  kind=params["kind"], precision=precision,
                                  deformation_width=params["deformation_width"],
                                  attachment_width=params["attachment_width"],
                                  data_sigma_squared=params["data_sigma_squared"],
                                  number_of_time_points=params["T"],
                                  improved_euler=params["improved_euler"], checkpoint_every=3,
                                  adjoint="continuous")
    return HIPAtlasModel(atlas, fixed, freeze_template=freeze_template,
                         smoothing_width=smoothing)


def _amygdala_gpu(precision="fp64"):
    from deformetrica_code_amd.estimators import FastGradientAscent
    V0, F, targets = fc.amygdala()
    fixed, ind = fc.initial_effects(V0, F, 10.0, 1)
    model = _model(V0, F, targets, fc.AMYG_PARAMS, fixed, freeze_template=True, smoothing=0.0,
                   precision=precision)
    return FastGradientAscent(model, **fc.AMYG_ESTIMATOR).Update({}, ind)


def _banana_gpu(precision="fp64"):
    from deformetrica_code_amd.estimators import GradientAscent
    V0, F, targets = fc.banana()
    fixed, ind = fc.initial_effects(V0, F, 0.3, 3)
    model = _model(V0, F, targets, fc.BANANA_PARAMS, fixed, freeze_template=False,
                   smoothing=fc.BANANA_SMOOTHING, precision=precision)
    return GradientAscent(model, **fc.BANANA_ESTIMATOR).Update({}, ind)


def _worst_difference(out, ref):
    """Largest relative difference over every floating field of two snapshots."""
    worst = 0.0
    for a, b in zip(out["log_likelihood_history"], ref["log_likelihood_history"]):
        if b.size:
            worst = max(worst, rel(a, b))
    for key in ("fixed_effects", "aux_fixed_effects", "pop_grad", "individual_rer",
                "aux_individual_rer", "ind_grad"):
        if key not in out:
            continue
        for name, v in out[key].items():
            a = np.concatenate([np.ravel(x) for x in v]) if isinstance(v, list) else v
            w = ref[key][name]
            b = np.concatenate([np.ravel(x) for x in w]) if isinstance(w, list) else w
            worst = max(worst, rel(a, b))
    if "grad_sq_norms" in out:
        worst = max(worst, rel(out["grad_sq_norms"], ref["grad_sq_norms"]))
    return worst


def test_gpu_loglik_at_the_stored_amygdala_state():
    from deformetrica_code_amd.atlas import HIPDeterministicAtlas
    from oracle import atlas as oa
    V0, F, targets = fc.amygdala()
    p = fc.AMYG_PARAMS
    atlas = HIPDeterministicAtlas(F, V0.shape[0], 8, targets, kind="current",
                                  deformation_width=p["deformation_width"],
                                  attachment_width=p["attachment_width"],
                                  data_sigma_squared=p["data_sigma_squared"],
                                  number_of_time_points=p["T"], improved_euler=False)
    f = fc.snapshot(fc.AMYG_STATE)
    cp = f["fixed_effects"]["ControlPoints"]
    mom = np.stack(f["individual_rer"]["Momenta"])
    (data, reg), _ = atlas.ComputeLogLikelihood(cp, mom, V0)
    stored = f["log_likelihood_history"][10]
    assert abs(reg - stored[1]) <= 1e-12 * abs(stored[1])
    (data0, reg0), _ = atlas.ComputeLogLikelihood(cp, np.zeros_like(mom), V0)
    d_or, _, _ = oa.loglik(cp, np.zeros_like(mom), V0, F, targets, **p)
    stored0 = f["log_likelihood_history"][0][0]
    assert reg0 == 0.0
    assert abs(data0 - d_or) <= 1e-12 * abs(d_or)
    assert abs(data0 - stored0) <= 1e-12 * abs(stored0)
    print(f"data term at the undeformed template: GPU {data0!r}, oracle {d_or!r}, "
          f"stored {stored0!r}")


def test_gpu_replay_of_the_banana_atlas():
    from deformetrica_code_amd.io import states_close
    print("GPU GradientAscent replay of the banana atlas, fp64, continuous adjoint", flush=True)
    out = _banana_gpu()
    ref = fc.snapshot(fc.BANANA_STATE)
    assert out["iter"] == ref["iter"] == 9
    np.testing.assert_array_equal(out["step"], ref["step"])
    assert states_close(fc.banana_state(out), fc.banana_state(ref), 1e-5)
    assert states_close(fc.banana_state(out), fc.banana_state(ref), 1e-10)
    worst = _worst_difference(out, fc.banana_oracle_replay())
    print(f"largest relative difference GPU vs oracle replay: {worst:.3e}")
    assert worst < 1e-9


def test_gpu_replay_of_the_amygdala_registration():
    print("GPU FastGradientAscent replay of the amygdala registration, fp64", flush=True)
    out = _amygdala_gpu()
    ref_oracle = fc.amygdala_oracle_replay()
    for k in ("iter", "freeze_direction_counter", "tau"):
        assert out[k] == ref_oracle[k]
    np.testing.assert_array_equal(out["step"], ref_oracle["step"])
    worst = _worst_difference(out, ref_oracle)
    print(f"largest relative difference GPU vs oracle replay: {worst:.3e}")
    assert worst < 1e-9
    fc.compare_amygdala_with_snapshot(out, fc.snapshot(fc.AMYG_STATE))


def test_gpu_replays_in_fp32c_take_the_same_paths():
    """FP32C kernel sums (FP64 at the boundary): the same accepted steps, values close to
    the FP64 replays."""
    for name, run in (("amygdala", _amygdala_gpu), ("banana", _banana_gpu)):
        print(f"GPU replay {name}, fp32c", flush=True)
        out, ref = run("fp32c"), run("fp64")
        assert out["iter"] == ref["iter"]
        np.testing.assert_array_equal(out["step"], ref["step"])
        d = max(rel(a, b) for a, b in zip(out["log_likelihood_history"],
                                          ref["log_likelihood_history"]) if b.size)
        dm = rel(np.concatenate(out["individual_rer"]["Momenta"]),
                 np.concatenate(ref["individual_rer"]["Momenta"]))
        print(f"{name} fp32c vs fp64: history {d:.3e}, momenta {dm:.3e}")
        assert d < 1e-5 and dm < 1e-5


def _multi_gpu(case, precision="fp64"):
    """FastGradientAscent over HIPMultiObjectAtlasModel for the skulls, parabola or (as a
    one-surface multi-object) amygdala test."""
    from deformetrica_code_amd.atlas import HIPMultiObjectAtlasModel
    from deformetrica_code_amd.curves import HIPNonOrientedPolyLine, HIPOrientedPolyLine
    from deformetrica_code_amd.estimators import FastGradientAscent
    from deformetrica_code_amd.surfaces import HIPNonOrientedSurfaceMesh, HIPOrientedSurfaceMesh
    if case == "amygdala":
        template, objects, targets = fc.amygdala_multi()
        p = fc.AMYG_PARAMS
        spacing, estimator, smoothing = 10.0, fc.AMYG_ESTIMATOR, 0.0
    else:
        template, objects, targets = getattr(fc, case)()
        p = getattr(fc, case.upper() + "_PARAMS")
        spacing, estimator = p["deformation_width"], getattr(fc, case.upper() + "_ESTIMATOR")
        smoothing = fc.PARABOLA_SMOOTHING if case == "parabola" else 0.0
    fixed, ind = fc.multi_initial_effects(template, objects, spacing, len(targets))

    def make(kind, points, cells, width):
        if np.asarray(cells).shape[1] == 3:
            cls = HIPOrientedSurfaceMesh if kind == "current" else HIPNonOrientedSurfaceMesh
        else:
            cls = HIPOrientedPolyLine if kind == "current" else HIPNonOrientedPolyLine
        return cls(points, cells, width, precision)

    tpl = [make(k, v, c, w) for v, (k, c, w, _) in zip(template, objects)]
    tg = [[make(k, v, c, w) for (v, c), (k, _, w, _) in zip(t, objects)] for t in targets]
    model = HIPMultiObjectAtlasModel(tpl, [o[3] for o in objects], tg, fixed,
                                     deformation_width=p["deformation_width"],
                                     number_of_time_points=p["T"],
                                     improved_euler=p["improved_euler"], precision=precision,
                                     freeze_template=case != "parabola",
                                     smoothing_width=smoothing, checkpoint_every=3)
    return FastGradientAscent(model, **estimator).Update({}, ind)


@pytest.mark.parametrize("case,tolerance", [("skulls", 1e-7), ("parabola", 1e-8)])
def test_gpu_replays_of_the_2d_polyline_tests(case, tolerance):
    from deformetrica_code_amd.io import DeformationState, push_fast_gradient_ascent, states_close
    print(f"GPU FastGradientAscent replay of {case}, fp64, multi-object model", flush=True)
    out = _multi_gpu(case)
    ref_oracle = getattr(fc, case + "_oracle_replay")()
    for k in ("iter", "freeze_direction_counter", "tau"):
        assert out[k] == ref_oracle[k]
    np.testing.assert_array_equal(out["step"], ref_oracle["step"])
    worst = _worst_difference(out, ref_oracle)
    print(f"largest relative difference GPU vs oracle replay: {worst:.3e}")
    assert worst < 1e-9
    ref = fc.snapshot(getattr(fc, case.upper() + "_STATE"))
    assert out["iter"] == ref["iter"] and out["tau"] == ref["tau"]
    np.testing.assert_array_equal(out["step"], ref["step"])
    mine = push_fast_gradient_ascent(DeformationState(), out)
    stored = push_fast_gradient_ascent(DeformationState(), ref)
    assert states_close(mine, stored, tolerance)
    assert states_close(mine, stored, 1e-10, rtol=1e-12)


def test_multi_object_model_agrees_with_the_batched_atlas():
    """The amygdala registration through HIPMultiObjectAtlasModel (one surface object,
    HIPDiffeos per subject) and through the batched C-ABI atlas: the same path, values
    within 1e-9 relative."""
    print("GPU amygdala replay through the multi-object model", flush=True)
    out, ref = _multi_gpu("amygdala"), _amygdala_gpu()
    assert out["iter"] == ref["iter"] and out["tau"] == ref["tau"]
    np.testing.assert_array_equal(out["step"], ref["step"])
    worst = _worst_difference(out, ref)
    print(f"largest relative difference multi-object vs batched atlas: {worst:.3e}")
    assert worst < 1e-9


@pytest.mark.parametrize("case", ["skulls", "parabola"])
def test_gpu_2d_replays_in_fp32c_take_the_same_paths(case):
    print(f"GPU replay {case}, fp32c", flush=True)
    out, ref = _multi_gpu(case, "fp32c"), _multi_gpu(case, "fp64")
    assert out["iter"] == ref["iter"]
    np.testing.assert_array_equal(out["step"], ref["step"])
    d = max(rel(a, b) for a, b in zip(out["log_likelihood_history"],
                                      ref["log_likelihood_history"]) if b.size)
    dm = rel(np.concatenate(out["individual_rer"]["Momenta"]),
             np.concatenate(ref["individual_rer"]["Momenta"]))
    print(f"{case} fp32c vs fp64: history {d:.3e}, momenta {dm:.3e}")
    assert d < 1e-5 and dm < 1e-5


def _regression_gpu(precision="fp64"):
    from deformetrica_code_amd.curves import HIPNonOrientedPolyLine
    from deformetrica_code_amd.estimators import FastGradientAscent
    from deformetrica_code_amd.regression import HIPRegressionModel
    template, objects, obs, ti = fc.regression_skulls()
    (kind, E, width, s2), = objects
    assert kind == "varifold"
    tpl = [HIPNonOrientedPolyLine(template[0], E, width, precision)]
    targets = [[HIPNonOrientedPolyLine(v, c, width, precision)] for (v, c), in obs]
    p = fc.REGRESSION_PARAMS
    model = HIPRegressionModel(tpl, [s2], targets, ti, fc.regression_initial_effects(template, objects),
                               deformation_width=p["deformation_width"], number_of_time_points=p["T"],
                               t0=p["t0"], tN=p["tN"], improved_euler=p["improved_euler"],
                               precision=precision, smoothing_width=fc.REGRESSION_SMOOTHING,
                               checkpoint_every=2)
    return FastGradientAscent(model, **fc.REGRESSION_ESTIMATOR).Update({}, {})


def test_gpu_replay_of_the_skulls_regression():
    """Test7 through HIPRegressionModel (the adjoint with jump conditions on the GPU): within
    1e-9 relative of the oracle replay, the stored snapshot's path exactly, every value
    within the test's 1e-8 and within 1e-10 plus 1e-12 of its magnitude."""
    from deformetrica_code_amd.io import DeformationState, push_fast_gradient_ascent, states_close
    print("GPU FastGradientAscent replay of the skulls regression, fp64", flush=True)
    out = _regression_gpu()
    ref_oracle = fc.regression_oracle_replay()
    for k in ("iter", "freeze_direction_counter", "tau"):
        assert out[k] == ref_oracle[k]
    np.testing.assert_array_equal(out["step"], ref_oracle["step"])
    worst = _worst_difference(out, ref_oracle)
    print(f"largest relative difference GPU vs oracle replay: {worst:.3e}")
    assert worst < 1e-9
    ref = fc.snapshot(fc.REGRESSION_STATE)
    np.testing.assert_array_equal(out["step"], ref["step"])
    mine = push_fast_gradient_ascent(DeformationState(), out)
    stored = push_fast_gradient_ascent(DeformationState(), ref)
    assert states_close(mine, stored, 1e-8)
    assert states_close(mine, stored, 1e-10, rtol=1e-12)


def test_gpu_regression_in_fp32c_takes_the_same_path():
    print("GPU regression replay, fp32c", flush=True)
    out, ref = _regression_gpu("fp32c"), _regression_gpu("fp64")
    assert out["iter"] == ref["iter"]
    np.testing.assert_array_equal(out["step"], ref["step"])
    d = max(rel(a, b) for a, b in zip(out["log_likelihood_history"],
                                      ref["log_likelihood_history"]) if b.size)
    dm = rel(out["fixed_effects"]["InitialMomenta"], ref["fixed_effects"]["InitialMomenta"])
    print(f"regression fp32c vs fp64: history {d:.3e}, momenta {dm:.3e}")
    assert d < 1e-5 and dm < 1e-5


def test_skulls_registration_through_the_batched_atlas():
    """Test5 through the batched C-ABI atlas (HIPDeterministicAtlas with segments in 2D,
    continuous adjoint) under HIPAtlasModel: the oracle replay's path, values within 1e-9."""
    from deformetrica_code_amd.atlas import HIPAtlasModel, HIPDeterministicAtlas
    from deformetrica_code_amd.estimators import FastGradientAscent
    template, objects, targets = fc.skulls()
    (kind, E, width, s2), = objects
    p = fc.SKULLS_PARAMS
    fixed, ind = fc.multi_initial_effects(template, objects, p["deformation_width"], 1)
    atlas = HIPDeterministicAtlas(E, len(template[0]), fixed["ControlPoints"].shape[0],
                                  [t[0] for t in targets], kind=kind,
                                  deformation_width=p["deformation_width"],
                                  attachment_width=width, data_sigma_squared=s2,
                                  number_of_time_points=p["T"],
                                  improved_euler=p["improved_euler"], checkpoint_every=2,
                                  adjoint="continuous")
    model = HIPAtlasModel(atlas, fixed, freeze_template=True)
    print("skulls registration through the batched atlas", flush=True)
    out = FastGradientAscent(model, **fc.SKULLS_ESTIMATOR).Update({}, ind)
    ref = fc.skulls_oracle_replay()
    for k in ("iter", "freeze_direction_counter", "tau"):
        assert out[k] == ref[k]
    np.testing.assert_array_equal(out["step"], ref["step"])
    worst = _worst_difference(out, ref)
    print(f"largest relative difference batched atlas vs oracle replay: {worst:.3e}")
    assert worst < 1e-9
