"""The reference's amygdala registration (configure.ini Test6) run end to end on the GPU:
FastGradientAscent (deformetrica_code_amd.estimators) over the batched atlas with one
subject, a frozen template and the reference's continuous adjoint, in FP64.

* against the same estimator over the oracle model: every field of the final snapshot at
  1e-9 relative (ten iterations of an optimizer over GPU sums in another order);
* against the reference's stored snapshot: the discrete path exactly and the values at the
  level set by the stored data term's offset (see tests/test_io.py);
* the GPU log-likelihood at the stored state: the regularity term to 1e-12 of the stored
  one, the data term of the undeformed template to 1e-12 of the oracle's.
"""
import numpy as np
import pytest

from amygdala_case import (AMYG_PARAMS, OracleRegistration, amygdala, amygdala_initial_effects,
                           compare_with_snapshot, snapshot)

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _gpu_model(precision="fp64"):
    from deformetrica_code_amd.atlas import HIPAtlasModel, HIPDeterministicAtlas
    V0, F, Vt, Ft = amygdala()
    p = AMYG_PARAMS
    atlas = HIPDeterministicAtlas(F, V0.shape[0], 8, [(Vt, Ft)], kind=p["kind"],
                                  precision=precision,
                                  deformation_width=p["deformation_width"],
                                  attachment_width=p["attachment_width"],
                                  data_sigma_squared=p["data_sigma_squared"],
                                  number_of_time_points=p["T"], checkpoint_every=3,
                                  adjoint="continuous")
    return HIPAtlasModel(atlas, freeze_template=True), atlas


def _replay(model):
    from deformetrica_code_amd.estimators import FastGradientAscent
    fixed, ind = amygdala_initial_effects()
    est = FastGradientAscent(model, max_iterations=10, initial_step_size=1e-7)
    return est.Update(fixed, {}, ind)


def test_gpu_loglik_at_the_stored_state():
    from oracle import atlas as oa
    _, atlas = _gpu_model()
    V0, F, Vt, Ft = amygdala()
    f = snapshot()
    cp = f["fixed_effects"]["ControlPoints"]
    mom = np.stack(f["individual_rer"]["Momenta"])
    (data, reg), _ = atlas.ComputeLogLikelihood(cp, mom, V0)
    stored = f["log_likelihood_history"][10]
    assert abs(reg - stored[1]) <= 1e-12 * abs(stored[1])
    (data0, reg0), _ = atlas.ComputeLogLikelihood(cp, np.zeros_like(mom), V0)
    d_or, _, _ = oa.loglik(cp, np.zeros_like(mom), V0, F, [(Vt, Ft)], **AMYG_PARAMS)
    assert reg0 == 0.0
    assert abs(data0 - d_or) <= 1e-12 * abs(d_or)
    print(f"data term at the undeformed template: GPU {data0!r}, oracle {d_or!r}, "
          f"stored {f['log_likelihood_history'][0][0]!r}")


def test_gpu_replay_of_the_amygdala_registration():
    print("GPU FastGradientAscent replay, fp64, continuous adjoint", flush=True)
    model, _ = _gpu_model()
    out = _replay(model)
    ref_oracle = _replay(OracleRegistration())
    for k in ("iter", "freeze_direction_counter", "tau"):
        assert out[k] == ref_oracle[k]
    np.testing.assert_array_equal(out["step"], ref_oracle["step"])
    worst = 0.0
    for a, b in zip(out["log_likelihood_history"], ref_oracle["log_likelihood_history"]):
        worst = max(worst, rel(a, b))
    for key, sub in (("fixed_effects", "ControlPoints"), ("aux_fixed_effects", "ControlPoints"),
                     ("pop_grad", "ControlPoints")):
        worst = max(worst, rel(out[key][sub], ref_oracle[key][sub]))
    for key in ("individual_rer", "aux_individual_rer", "ind_grad"):
        worst = max(worst, rel(out[key]["Momenta"][0], ref_oracle[key]["Momenta"][0]))
    worst = max(worst, rel(out["grad_sq_norms"], ref_oracle["grad_sq_norms"]))
    print(f"largest relative difference GPU vs oracle replay: {worst:.3e}")
    assert worst < 1e-9
    compare_with_snapshot(out, snapshot())


def test_gpu_replay_fp32c_takes_the_same_path():
    """FP32C kernel sums (FP64 at the boundary): the same accepted steps, values close to
    the FP64 replay."""
    print("GPU FastGradientAscent replay, fp32c", flush=True)
    out = _replay(_gpu_model("fp32c")[0])
    ref = _replay(_gpu_model("fp64")[0])
    for k in ("iter", "freeze_direction_counter", "tau"):
        assert out[k] == ref[k]
    np.testing.assert_array_equal(out["step"], ref["step"])
    d = max(rel(a, b) for a, b in zip(out["log_likelihood_history"], ref["log_likelihood_history"]))
    dm = rel(out["individual_rer"]["Momenta"][0], ref["individual_rer"]["Momenta"][0])
    print(f"fp32c vs fp64: history {d:.3e}, momenta {dm:.3e}")
    assert d < 1e-5 and dm < 1e-6  # measured 3.5e-7 and 2.1e-8
