"""Geodesic regression on the GPU: Deformetrica's Regression model (Regression.cxx) for a
template of landmark objects (polylines, triangle surfaces, volume meshes, point clouds and
landmarks), as the estimators see it.

One geodesic is shot from the control points and the initial momenta; the template flows
along it and each observation is matched at its time index of the time grid.  Every
kernel sum runs on the GPU: the shooting, the flow and the reference's continuous adjoint
with the observations as jump conditions through HIPDiffeos (dfm_shoot_flow_adjoint_jumps),
the data terms through the objects' own ComputeMatch / ComputeMatchGradient, the RKHS term
and the template smoothing through HIPExactKernel.
"""
from __future__ import annotations

import numpy as np

from .atlas import _copy_effects


def time_index(age, T, t0, tN):
    """The time index of an observation at ``age`` (estimate_geodesic_regression.cxx:
    229-234): int((age - t0) * ((T - 1) / (tN - t0)) + 0.5f)."""
    return int((age - t0) * ((T - 1) / (tN - t0)) + np.float32(0.5))


class HIPRegressionModel:
    """Regression (Regression.cxx) over GPU kernel sums.

    ``objects`` are the template's data-term objects (kernel widths set),
    ``data_sigma_squared`` one noise variance per object, ``observations`` one list of
    target objects (in object order) per observation, ``time_indices`` their indices on the
    grid of ``number_of_time_points`` points over [t0, tN].  Fixed effects "ControlPoints"
    (n x d), "InitialMomenta" (n x d) and "TemplateData" (one point array per object); no
    individual effects.

    The terms follow the reference:
    - log-likelihood [-1/2 sum_j sum_i r_ji / data_sigma_squared_i, -1/2 m^T K(cp, cp) m]
      (UpdateFixedEffectsAndComputeCompleteLogLikelihood, :198-234), r_ji the match of
      object i of the template flowed to observation j's time index;
    - gradient (ComputeDataTermGradient, :600-662): each observation's match gradients
      divided by 2 data_sigma_squared_i, the continuous adjoint with those conditions at
      the time indices, then -xi_pos(0), -xi_mom(0) and -theta(0), the last smoothed per
      object with the "TemplateData" fixed effect as sources (:679-705);
    - no regularity gradient: AddGradientRegularityTerm (:664-677) takes the gradients by
      value, so the reference's gradient never holds that term;
    - SetFixedEffects (:54-66) moves the template to the new TemplateData at once (no lag),
      or keeps the old TemplateData when the template is frozen;
    - gradient blocks "ControlPoints" unless frozen, "InitialMomenta", "TemplateData"
      unless frozen, squared norms in that order (:306-316).
    The bounding-box test of the reference's Diffeos is not made: out_of_box is always
    False."""

    def __init__(self, objects, data_sigma_squared, observations, time_indices, fixed_effects,
                 *, deformation_width, number_of_time_points, t0=0.0, tN=1.0,
                 improved_euler=False, scheme="euler", precision="fp64", freeze_template=False,
                 freeze_control_points=False, smoothing_width=0.0, checkpoint_every=0):
        if len(data_sigma_squared) != len(objects):
            raise ValueError("one data sigma squared per object")
        if len(observations) != len(time_indices):
            raise RuntimeError("Mismatch between number of observations and number of time "
                               "indices.")
        if any(len(o) != len(objects) for o in observations):
            raise ValueError("every observation needs one target per template object")
        T = int(number_of_time_points)
        if any(not 0 <= int(t) < T for t in time_indices):
            raise ValueError("time indices must lie in [0, number_of_time_points)")
        self.objects = list(objects)
        self.data_sigma_squared = [float(s) for s in data_sigma_squared]
        self.observations = [list(o) for o in observations]
        self.time_indices = [int(t) for t in time_indices]
        self.deformation_width = float(deformation_width)
        self.T = T
        self.t0, self.tN = float(t0), float(tN)
        self.improved_euler = bool(improved_euler)
        self.scheme = scheme
        self.precision = precision
        self.freeze_template = bool(freeze_template)
        self.freeze_control_points = bool(freeze_control_points)
        self.smoothing_width = float(smoothing_width)
        self.checkpoint_every = int(checkpoint_every)
        self._fixed = _copy_effects(fixed_effects)
        self._counts = [o.GetNumberOfPoints() for o in self.objects]
        if [len(v) for v in self._fixed["TemplateData"]] != self._counts:
            raise ValueError("TemplateData must hold each object's points")
        self.dim = self._fixed["ControlPoints"].shape[1]

    def GetFixedEffects(self):
        return _copy_effects(self._fixed)

    def SetFixedEffects(self, fixed):
        template_memory = self._fixed["TemplateData"]
        self._fixed = _copy_effects(fixed)
        if self.freeze_template:
            self._fixed["TemplateData"] = template_memory

    def _diffeos(self):
        from .diffeos import HIPDiffeos
        d = HIPDiffeos(self.dim, self.precision)
        d.SetKernelWidth(self.deformation_width)
        d.SetNumberOfTimePoints(self.T)
        d.SetT0(self.t0)
        d.SetTN(self.tN)
        d.SetShootingScheme(self.scheme)
        if self.improved_euler:
            d.UseImprovedEuler()
        else:
            d.UseStandardEuler()
        d.SetStartPositions(self._fixed["ControlPoints"])
        d.SetStartMomentas(self._fixed["InitialMomenta"])
        d.SetLandmarkPoints(np.concatenate(self._fixed["TemplateData"]))
        d.Update()
        return d

    def _deformed_at(self, X, t):
        parts = np.split(X[t], np.cumsum(self._counts)[:-1])
        return [o.DeformedObject(x) for o, x in zip(self.objects, parts)]

    def UpdateFixedEffectsAndComputeCompleteLogLikelihood(self, pop_rer, ind_rer):
        X = self._diffeos().GetLandmarkPointsTrajectory()
        data = 0.0
        for obs, t in zip(self.observations, self.time_indices):
            for obj, tgt, s2 in zip(self._deformed_at(X, t), obs, self.data_sigma_squared):
                data -= 0.5 * obj.ComputeMatch(tgt) / s2
        from .kernel import HIPExactKernel
        cp, mom = self._fixed["ControlPoints"], self._fixed["InitialMomenta"]
        k = HIPExactKernel(self.dim, self.deformation_width, precision=self.precision)
        k.SetSources(cp)
        k.SetWeights(mom)
        reg = 0.0
        reg -= float(np.sum(k.Convolve(cp) * mom))
        reg *= 0.5
        return np.array([data, reg]), False

    def _smooth(self, gl2):
        if self.smoothing_width < 1e-20:
            return gl2
        from .kernel import HIPExactKernel
        out = []
        for tpl, g in zip(self._fixed["TemplateData"], gl2):
            k = HIPExactKernel(self.dim, self.smoothing_width, precision=self.precision)
            k.SetSources(tpl)
            k.SetWeights(g)
            out.append(k.Convolve(tpl))
        return out

    def ComputeCompleteLogLikelihoodGradient(self, pop_rer, ind_rer):
        d = self._diffeos()
        X = d.GetLandmarkPointsTrajectory()
        conditions = [np.concatenate([obj.ComputeMatchGradient(tgt) / (2.0 * s2) for obj, tgt, s2
                                      in zip(self._deformed_at(X, t), obs,
                                             self.data_sigma_squared)])
                      for obs, t in zip(self.observations, self.time_indices)]
        xp, xm, th0 = d.IntegrateAdjointEquations(conditions, self.checkpoint_every,
                                                  jump_times=self.time_indices)
        gcp, gmom = -xp, -xm
        pop, sq = {}, []
        if not self.freeze_control_points:
            pop["ControlPoints"] = gcp
            sq.append(float(np.sum(gcp * gcp)))
        pop["InitialMomenta"] = gmom
        sq.append(float(np.sum(gmom * gmom)))
        if not self.freeze_template:
            gl2 = np.split(-th0, np.cumsum(self._counts)[:-1])
            sob = self._smooth(gl2)
            pop["TemplateData"] = sob
            sq.append(sum(float(np.sum(a * b)) for a, b in zip(gl2, sob)))
        return pop, {}, np.asarray(sq)


__all__ = ["HIPRegressionModel", "time_index"]
